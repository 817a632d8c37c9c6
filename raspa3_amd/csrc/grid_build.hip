// Tricubic interpolation-grid builder for gfx950.
//
// Replaces InterpolationEnergyGrid::makeFrameworkInterpolationGrid (tricubic branch,
// src/raspakit/interpolation_energy_grid.cpp:3800-3875) and the per-point analytic sum
// Interactions::calculateTricubicFractionalAtPosition (src/raspakit/interactions/framework_molecule_grid.cpp:168-282).
//
// The reference visits every supercell atom at every point and applies the minimum-image convention per pair.
// With a supercell at least twice the VDW cutoff wide, the set of in-cutoff pairs it forms is exactly the set of
// periodic images of the unit-cell atoms within the cutoff of the point, so this builder enumerates those images
// once per framework, bins them in fractional space, and lets each workgroup gather only the images that can be
// within the cutoff of its tile of grid points.
//
// Work decomposition: one workgroup = 256 lanes = an 8x8x4 tile of grid points, each wavefront a 4x4x4
// sub-block, one grid point per lane.  Candidate images are staged through LDS in chunks of 256; every lane of a
// wavefront reads the same candidate (an LDS broadcast), so the per-candidate branches (cutoff, LJ-active,
// charged) are uniform or nearly so.  All arithmetic is FP64.  Several grids of equal dimensions (LJ grids for
// different probe pseudo-atoms and the unit-charge Ewald real-space grid) are built in one pass and share the
// distance work.
//
// Triclinic fractional derivatives use the lattice vectors h_a and the metric G_ab = h_a.h_b instead of the
// reference's full 3x3x3 Cartesian tensors: with u_a = h_a.dr,
//   dE/ds_a          = sum f1 u_a
//   d2E/ds_a ds_b    = sum f2 u_a u_b + G_ab sum f1
//   d3E/ds0 ds1 ds2  = sum f3 u0 u1 u2 + G01 sum f2 u2 + G02 sum f2 u1 + G12 sum f2 u0
// which equals the reference's contraction (framework_molecule_grid.cpp:244-278) term by term.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ewald_tables.hpp"
#include "quintic_terms.hpp"
#include "mc_device.hpp"
#include "rg_internal.hpp"

namespace
{

struct alignas(16) Img
{
  double x, y, z, q;    // Cartesian position of the image, charge (e)
  double u0, u1, u2;    // u_a = h_a . x (lattice-vector projections, for triclinic derivatives)
  int type;             // local framework type row
  float qf;             // charge times the Coulomb factor in FP32 (filled when a launch stages the image)
};
static_assert(sizeof(Img) == 64, "Img must be 64 bytes");

struct alignas(16) LJPar
{
  double e4, s2, sh, c1, c2, c3, p0, p1;  // 4eps, sigma^2, shift, -6*4eps, 24*4eps, -96*4eps
};

constexpr int TI = 8, TJ = 8, TK = 4;
constexpr int MAXG = 4;    // LJ grids per launch
constexpr int MAXT = 16;   // distinct framework pseudo-atom types
constexpr int CH = 256;    // candidates per LDS chunk
constexpr int MAXROWS = 1024;
constexpr int kSplitDeg = 6;  // degree of the FP32 erfc polynomial of the split-precision Ewald grid

struct GridArgs
{
  double cell[9];
  double delta[3];
  double ncm1[3];
  double G00, G11, G22, G01, G02, G12;
  double diag[3];
  double rc2;
  double cut2;
  double ext[3];
  double lo[3], bw[3];
  int nb[3];
  int n[3];
  int ntile[3];
  int ntiles;
  const Img* img;
  const int* bin_start;
  const LJPar* ljpar;  // [NLJ][MAXT]
  double alpha, alpha2, k2a;
  double coulomb;
  double overlap;
  double* out[MAXG + 1];
  // split precision of the Ewald real-space grid: images at least sqrt(rsplit2) from every point of a wavefront take
  // FP32 radial factors (x erfcx(x) / alpha at x = alpha r from a polynomial in t = xs / r + xo, |t| <= 1)
  double rsplit2;
  double a2l2e;  // alpha^2 / ln 2
  float xs32, xo32, k2a32, alpha2_32;
  float ecx[kSplitDeg + 1];
};

// The hardware reciprocal and reciprocal square root are within 5.2e-8 relative (measured over 1e6 arguments in
// [0.1, 144]); one third-order correction brings them to FP64 rounding: with e the residual, 1/x = y (1 + e + e^2)
// and 1/sqrt(x) = y (1 + e/2 + 3e^2/8), truncation errors of order e^3 < 1e-21.
__device__ __forceinline__ double rcp64(double x)
{
  const double y = __builtin_amdgcn_rcp(x);
  const double e = fma(-x, y, 1.0);  // 1 - x y
  return fma(y, fma(e, e, e), y);
}

__device__ __forceinline__ double rsq64(double x)
{
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);  // 1 - x y^2
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// exp(-y) for y >= 0: Cody-Waite reduction exp(-y) = 2^-n exp(r) with n = rint(y / ln 2) and r = n ln 2 - y formed
// in two FMA steps (ln 2 split into its nearest double and the remainder), |r| <= ln(2)/2; exp(r) from the degree-11
// Chebyshev-derived polynomial of ewald_tables.hpp (within 2.3e-16 relative); ldexp underflows to 0 for huge y.
__device__ __forceinline__ double expNeg(double y)
{
  constexpr double kInvLn2 = 1.4426950408889634;
  constexpr double kLn2Hi = 0.69314718055994529;
  constexpr double kLn2Lo = 2.3190468138462996e-17;
  const double n = rint(y * kInvLn2);
  double r = fma(n, kLn2Hi, -y);
  r = fma(n, kLn2Lo, r);
  // Estrin's scheme: pairs c_2k + c_2k+1 r, combined with r^2, then r^4
  static_assert(rg::kExpDeg == 11, "Estrin tree below is written for degree 11");
  const double* c = rg::kExpCoef;
  const double r2 = r * r, r4 = r2 * r2;
  const double p01 = fma(c[1], r, c[0]), p23 = fma(c[3], r, c[2]), p45 = fma(c[5], r, c[4]);
  const double p67 = fma(c[7], r, c[6]), p89 = fma(c[9], r, c[8]), pab = fma(c[11], r, c[10]);
  const double p03 = fma(p23, r2, p01), p47 = fma(p67, r2, p45), p8b = fma(pab, r2, p89);
  const double p = fma(fma(p8b, r4, p47), r4, p03);
  return ldexp(p, -static_cast<int>(n));
}

// LDS image of the erfcx tables of ewald_tables.hpp: the fine pieces (10 coefficients each), then per tail piece
// {midpoint, inverse half-width, 17 coefficients}
constexpr int kFineW = rg::kErfcxFineDeg + 1;
constexpr int kTailW = rg::kErfcxTailDeg + 3;
constexpr int kErfcxTabLen = rg::kErfcxFinePieces * kFineW + rg::kErfcxTailPieces * kTailW;

__device__ __forceinline__ double erfcxTabEntry(int q)
{
  if (q < rg::kErfcxFinePieces * kFineW) return rg::kErfcxFine[q / kFineW][q % kFineW];
  q -= rg::kErfcxFinePieces * kFineW;
  const int piece = q / kTailW, k = q % kTailW;
  return k == 0 ? rg::kErfcxTailMid[piece] : (k == 1 ? rg::kErfcxTailInvHalf[piece] : rg::kErfcxTail[piece][k - 2]);
}

// erfcx(x) = exp(x^2) erfc(x) for x >= 0 from the piecewise Chebyshev interpolants of ewald_tables.hpp (within
// 2.3e-16 relative).  On [0, 4) every lane reads the coefficients of its own piece p = floor(8x) with
// t = 16x - (2p + 1), which is exact.  From 4 on (alpha r >= 4 inside the cutoff, rare) a wave-uniform branch
// evaluates the tail pieces; from 16 on the result is 0, where erfc(x) and exp(-x^2) are below 1e-111 and every
// real-space term vanishes to FP64 precision.
__device__ __forceinline__ double erfcxLds(double x, const double* __restrict__ tab)
{
  const int p = min(static_cast<int>(8.0 * x), rg::kErfcxFinePieces - 1);
  const double t = fma(16.0, x, -static_cast<double>(2 * p + 1));
  const double* c = tab + p * kFineW;
  double acc = c[rg::kErfcxFineDeg];
#pragma unroll
  for (int j = rg::kErfcxFineDeg - 1; j >= 0; --j) acc = fma(acc, t, c[j]);
  if (__ballot(x >= 4.0))
  {
    if (x >= 4.0)
    {
      const int q = x < 6.0 ? 0 : (x < 8.0 ? 1 : (x < 12.0 ? 2 : 3));
      const double* rec = tab + rg::kErfcxFinePieces * kFineW + q * kTailW;
      const double tt = (x - rec[0]) * rec[1];
      double a = rec[2 + rg::kErfcxTailDeg];
#pragma unroll
      for (int j = rg::kErfcxTailDeg - 1; j >= 0; --j) a = fma(a, tt, rec[2 + j]);
      acc = x < rg::kErfcxMax ? a : 0.0;
    }
  }
  return acc;
}

template <bool TRIC>
struct Acc
{
  double e, d1[3], d2[3], d3;  // d2: 01, 02, 12
  double s1, b[3];             // triclinic metric terms: sum f1, sum f2 u_a
  __device__ void zero()
  {
    e = d3 = s1 = 0.0;
    for (int a = 0; a < 3; ++a) d1[a] = d2[a] = b[a] = 0.0;
  }
  __device__ __forceinline__ void add(double ee, double f1, double f2, double f3, double u0, double u1, double u2,
                                      double u01, double u02, double u12, double u012)
  {
    e += ee;
    d1[0] = fma(f1, u0, d1[0]);
    d1[1] = fma(f1, u1, d1[1]);
    d1[2] = fma(f1, u2, d1[2]);
    d2[0] = fma(f2, u01, d2[0]);
    d2[1] = fma(f2, u02, d2[1]);
    d2[2] = fma(f2, u12, d2[2]);
    d3 = fma(f3, u012, d3);
    if (TRIC)
    {
      s1 += f1;
      b[0] = fma(f2, u0, b[0]);
      b[1] = fma(f2, u1, b[1]);
      b[2] = fma(f2, u2, b[2]);
    }
  }
};

// ---- tile set-up and candidate staging, shared by the tricubic and triquintic kernels

// The lane's grid point and its tile.
struct TilePoint
{
  int gi, gj, gk;
  bool active;
  double px, py, pz;     // Cartesian position of the point
  double pu0, pu1, pu2;  // h_a . p (triclinic only)
  double cxc, cyc, czc;  // centre of the tile's points
  int nrow;              // rows of bins in the tile's bin box
};

// Finds the lane's point and fills the tile's bin-row table: rowStart[r] = first image of row r, rowPrefix[r] =
// images in rows before r (rowPrefix[nrow] = total).  Ends with a barrier.
template <bool TRIC>
__device__ __forceinline__ TilePoint tileSetup(const GridArgs& A, int* rowStart, int* rowPrefix)
{
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  TilePoint T;

  // XCD-aware tile order: hardware places workgroup b on XCD b % 8; give each XCD a contiguous range of tiles.
  int bid = blockIdx.x;
  {
    const int nfull = (A.ntiles / 8) * 8;
    if (bid < nfull) bid = (bid % 8) * (A.ntiles / 8) + bid / 8;
  }
  const int tx = bid % A.ntile[0];
  const int ty = (bid / A.ntile[0]) % A.ntile[1];
  const int tz = bid / (A.ntile[0] * A.ntile[1]);

  T.gi = tx * TI + (wv & 1) * 4 + (lane & 3);
  T.gj = ty * TJ + (wv >> 1) * 4 + ((lane >> 2) & 3);
  T.gk = tz * TK + (lane >> 4);
  T.active = T.gi < A.n[0] && T.gj < A.n[1] && T.gk < A.n[2];
  // s = (i*delta_x, j*delta_y, k*delta_z); pos = cell * s (src/raspakit/interpolation_energy_grid.cpp:3829-3831)
  const double s0 = double(T.gi) * A.delta[0], s1 = double(T.gj) * A.delta[1], s2 = double(T.gk) * A.delta[2];
  T.px = A.cell[0] * s0 + A.cell[3] * s1 + A.cell[6] * s2;
  T.py = A.cell[1] * s0 + A.cell[4] * s1 + A.cell[7] * s2;
  T.pz = A.cell[2] * s0 + A.cell[5] * s1 + A.cell[8] * s2;
  T.pu0 = T.pu1 = T.pu2 = 0.0;
  if (TRIC)
  {
    T.pu0 = A.cell[0] * T.px + A.cell[1] * T.py + A.cell[2] * T.pz;
    T.pu1 = A.cell[3] * T.px + A.cell[4] * T.py + A.cell[5] * T.pz;
    T.pu2 = A.cell[6] * T.px + A.cell[7] * T.py + A.cell[8] * T.pz;
  }

  // centre of the tile's points (fractional and Cartesian; a tile at the grid edge holds fewer points) and the
  // bin box that holds every image within rc of them
  const double fc0 = 0.5 * double(2 * tx * TI + min(TI, A.n[0] - tx * TI) - 1) * A.delta[0];
  const double fc1 = 0.5 * double(2 * ty * TJ + min(TJ, A.n[1] - ty * TJ) - 1) * A.delta[1];
  const double fc2 = 0.5 * double(2 * tz * TK + min(TK, A.n[2] - tz * TK) - 1) * A.delta[2];
  T.cxc = A.cell[0] * fc0 + A.cell[3] * fc1 + A.cell[6] * fc2;
  T.cyc = A.cell[1] * fc0 + A.cell[4] * fc1 + A.cell[7] * fc2;
  T.czc = A.cell[2] * fc0 + A.cell[5] * fc1 + A.cell[8] * fc2;
  int blo[3], bhi[3];
  {
    const double fc[3] = {fc0, fc1, fc2};
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
      int l = int(floor((fc[a] - A.ext[a] - A.lo[a]) / A.bw[a]));
      int h = int(floor((fc[a] + A.ext[a] - A.lo[a]) / A.bw[a]));
      blo[a] = max(0, min(A.nb[a] - 1, l));
      bhi[a] = max(0, min(A.nb[a] - 1, h));
    }
  }
  const int n1r = bhi[1] - blo[1] + 1;
  const int nrow = n1r * (bhi[2] - blo[2] + 1);
  T.nrow = nrow;

  // rows of bins contiguous along axis 0: [start(blo0, by, bz), start(bhi0 + 1, by, bz))
  for (int r = tid; r < nrow; r += 256)
  {
    const int by = blo[1] + r % n1r;
    const int bz = blo[2] + r / n1r;
    const int base = A.nb[0] * (by + A.nb[1] * bz);
    const int st = A.bin_start[base + blo[0]];
    const int en = A.bin_start[base + bhi[0] + 1];
    rowStart[r] = st;
    rowPrefix[r + 1] = en - st;
  }
  if (tid == 0) rowPrefix[0] = 0;
  __syncthreads();
  if (wv == 0)
  {
    int carry = 0;
    for (int c0 = 0; c0 < nrow; c0 += 64)
    {
      const int r = c0 + lane;
      int v = (r < nrow) ? rowPrefix[r + 1] : 0;
      for (int o = 1; o < 64; o <<= 1)
      {
        int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
      }
      if (r < nrow) rowPrefix[r + 1] = v + carry;
      carry += __shfl(v, 63, 64);
    }
  }
  __syncthreads();
  return T;
}

// Gathers images [base, base + 256) of the tile's bin rows, keeps those within the tile's sphere that matter to the
// launch (charged for the Coulomb grid, Lennard-Jones-active for one of the NLJ probes), compacts them into cand
// with the charge scaled by the Coulomb factor, and returns their number.  Every thread of the workgroup calls it;
// it ends with a barrier.
template <int NLJ, bool COUL>
__device__ __forceinline__ int stageChunk(const GridArgs& A, const TilePoint& T, int base, int total,
                                          const int* rowStart, const int* rowPrefix, const LJPar (*par)[MAXT],
                                          Img* cand, int* waveCount)
{
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int g = base + tid;
  bool pass = false;
  double2 r0, r1, r2, r3;  // the image record as (x, y), (z, q), (u0, u1), (u2, type)
  if (g < total)
  {
    int lo = 0, hi = T.nrow - 1;
    while (lo < hi)
    {
      const int mid = (lo + hi + 1) >> 1;
      if (rowPrefix[mid] <= g)
        lo = mid;
      else
        hi = mid - 1;
    }
    const double2* src = reinterpret_cast<const double2*>(A.img + rowStart[lo] + (g - rowPrefix[lo]));
    r0 = src[0];
    r1 = src[1];
    r2 = src[2];
    r3 = src[3];
    const int type = static_cast<int>(__double_as_longlong(r3.y) & 0xffffffffll);
    const double dx = r0.x - T.cxc, dy = r0.y - T.cyc, dz = r1.x - T.czc;
    pass = dx * dx + dy * dy + dz * dz < A.cut2;
    bool useful = COUL && r1.y != 0.0;
#pragma unroll
    for (int q = 0; q < NLJ; ++q) useful = useful || par[q][type].e4 != 0.0;
    pass = pass && useful;
  }
  const unsigned long long m = __ballot(pass);
  if (lane == 0) waveCount[wv] = __popcll(m);
  __syncthreads();
  int off = 0;
  for (int q = 0; q < wv; ++q) off += waveCount[q];
  const int ncand = waveCount[0] + waveCount[1] + waveCount[2] + waveCount[3];
  if (pass)
  {
    r1.y *= A.coulomb;
    // the record's last word (type, qf): keep the type, store the scaled charge in FP32
    r3.y = __longlong_as_double(
        (__double_as_longlong(r3.y) & 0xffffffffll) |
        (static_cast<long long>(__float_as_uint(static_cast<float>(r1.y))) << 32));
    double2* dst = reinterpret_cast<double2*>(&cand[off + __popcll(m & lt_mask)]);
    dst[0] = r0;
    dst[1] = r1;
    dst[2] = r2;
    dst[3] = r3;
  }
  __syncthreads();
  return ncand;
}

template <int NLJ, bool COUL, bool TRIC, bool SPLIT = false>
__global__ __launch_bounds__(256) void k_grid_build(GridArgs A)
{
  constexpr int NG = NLJ + (COUL ? 1 : 0);
  __shared__ Img cand[CH];
  __shared__ LJPar par[NLJ > 0 ? NLJ : 1][MAXT];
  __shared__ int rowStart[MAXROWS];
  __shared__ int rowPrefix[MAXROWS + 1];
  __shared__ int waveCount[4];
  __shared__ __attribute__((aligned(16))) double erfcxTab[COUL ? kErfcxTabLen : 1];

  const int tid = threadIdx.x;
  if (NLJ > 0)
  {
    for (int q = tid; q < NLJ * MAXT; q += 256) (&par[0][0])[q] = A.ljpar[q];
  }
  if (COUL)
  {
    for (int q = tid; q < kErfcxTabLen; q += 256) erfcxTab[q] = erfcxTabEntry(q);
  }
  const TilePoint T = tileSetup<TRIC>(A, rowStart, rowPrefix);
  const int total = rowPrefix[T.nrow];
  const double px = T.px, py = T.py, pz = T.pz;

  Acc<TRIC> acc[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) acc[g].zero();
  // the split path's polynomial offset and leading coefficient, held in vector registers: an FMA reads at most one
  // scalar operand, and with both of a pair in scalar registers the compiler re-copied one into a vector register
  // for every pair
  float xo32 = A.xo32, lead32 = A.ecx[kSplitDeg];
  if (SPLIT)
  {
    asm volatile("" : "+v"(xo32));
    asm volatile("" : "+v"(lead32));
  }

  for (int base = 0; base < total; base += 256)
  {
    const int ncand = stageChunk<NLJ, COUL>(A, T, base, total, rowStart, rowPrefix, par, cand, waveCount);

    // ---- per-lane pair sums over the staged candidates (every lane reads the same record: an LDS broadcast).  The
    // four waves of a workgroup hide each other's LDS latency; a register prefetch of the next record cost four
    // moves per pair and ran slower.
    const double2* cb = reinterpret_cast<const double2*>(cand);
    for (int c = 0; c < ncand; ++c)
    {
      const double2 c0 = cb[4 * c], c1 = cb[4 * c + 1];  // (x, y), (z, q)
      const Img& b = cand[c];
      const double dx = px - c0.x, dy = py - c0.y, dz = pz - c1.x;
      const double rr = dx * dx + dy * dy + dz * dz;
      if (rr < A.rc2)
      {
        // src/raspakit/interactions/framework_molecule_grid.cpp:200; the FP32 path below takes rr itself, which the
        // host's rsplit^2 >= 0.1 keeps above the clamp
        const double rrc = fmax(rr, 0.1);
        double u0, u1, u2;
        float qf = 0.0f;  // the staged FP32 charge (triclinic: read with u2 in one 16-byte LDS read)
        if (TRIC)
        {
          const double2 c2 = cb[4 * c + 2], c3 = cb[4 * c + 3];  // (u0, u1), (u2, type | qf)
          u0 = T.pu0 - c2.x;
          u1 = T.pu1 - c2.y;
          u2 = T.pu2 - c3.x;
          qf = __uint_as_float(static_cast<unsigned>(static_cast<unsigned long long>(__double_as_longlong(c3.y)) >> 32));
          // formed here, qf keeps the record's last 16 bytes one LDS read; left to the FP32 branch, the compiler
          // split off a 4-byte read (and its address copy) there
          if (SPLIT) asm volatile("" : "+v"(qf));
        }
        else
        {
          u0 = dx;
          u1 = dy;
          u2 = dz;
        }
        const double u01 = u0 * u1, u02 = u0 * u2, u12 = u1 * u2, u012 = u01 * u2;
        if (NLJ > 0)
        {
          const double ir2 = rcp64(rrc);
          const double ir4 = ir2 * ir2, ir6 = ir4 * ir2;
#pragma unroll
          for (int q = 0; q < NLJ; ++q)
          {
            // src/raspakit/potentials/interpolation/tricubic_derivative_lj.ixx:33-46.  No test of e4 != 0: a probe
            // that does not interact with this type (e4 = 0, so c1 = c2 = c3 = 0 and the shift U(rc) = 0) adds exact
            // zeros, and the staging keeps only images that interact with some probe.
            const LJPar& P = par[q][b.type];
            const double x = P.s2 * ir2;
            const double t3 = x * x * x;
            const double t6 = t3 * t3;
            const double e = fma(P.e4, t6 - t3, -P.sh);
            const double f1 = P.c1 * fma(2.0, t6, -t3) * ir2;
            const double f2 = P.c2 * fma(7.0, t6, -2.0 * t3) * ir4;
            const double f3 = P.c3 * fma(28.0, t6, -5.0 * t3) * ir6;
            acc[q].add(e, f1, f2, f3, u0, u1, u2, u01, u02, u12, u012);
          }
        }
        // a Coulomb-only launch stages charged images only.  Both precisions only form the radial factors; the one
        // accumulation after the join keeps the accumulators in place (accumulating inside each branch made the
        // compiler copy all twelve of them out and back around the join, 24 moves per pair).
        if (COUL && (NLJ == 0 || c1.y != 0.0))
        {
          double e, f1, f2, f3;
          if (SPLIT && !__ballot(rr < A.rsplit2))
          {
            // every lane of the wavefront is at least the split radius from this image: the radial factors in FP32
            // (hardware rsq and exp2, x erfcx(x) / alpha at x = alpha r from the polynomial in 1/r).  The exponent
            // -alpha^2 r^2 / ln 2 = k + f (k = its floor) is reduced in FP64 so that only f in [0, 1) passes
            // through FP32.  With ecr = q erfc(alpha r) r and E = q 2 alpha / sqrt(pi) r^2 exp(-alpha^2 r^2), the
            // factors of the FP64 branch below become e = ecr / r^2, f1 = -(E + ecr) / r^4,
            // f2 = (2 E alpha^2 r^2 + 3 (E + ecr)) / r^6, f3 = -(E alpha^2 r^2 (4 alpha^2 r^2 + 10) + 15 (E + ecr)) / r^8.
            const float rr32 = static_cast<float>(rr);
            const float ir = __builtin_amdgcn_rsqf(rr32);
            const float t = fmaf(ir, A.xs32, xo32);
            float pe = lead32;
#pragma unroll
            for (int j = kSplitDeg - 1; j >= 0; --j) pe = fmaf(pe, t, A.ecx[j]);
            const double y = rr * -A.a2l2e;
            const double k = floor(y);
            const float ex = __builtin_amdgcn_ldexpf(__builtin_amdgcn_exp2f(static_cast<float>(y - k)),
                                                     static_cast<int>(k));
            const float qex = (TRIC ? qf : b.qf) * ex;         // q exp(-alpha^2 r^2)
            const float ecr = qex * pe;                        // q erfc(alpha r) r
            const float E = A.k2a32 * qex * rr32;
            const float ar2 = A.alpha2_32 * rr32;
            const float s = E + ecr;
            const float ear = E * ar2;
            const float ir2 = ir * ir;
            const float ir4 = ir2 * ir2;
            const float ir6 = ir4 * ir2;
            const float ir8 = ir4 * ir4;
            e = double(ecr * ir2);
            f1 = double(-s * ir4);
            f2 = double(fmaf(2.0f, ear, 3.0f * s) * ir6);
            f3 = double(-fmaf(ear, fmaf(4.0f, ar2, 10.0f), 15.0f * s) * ir8);
          }
          else
          {
            // src/raspakit/potentials/interpolation/tricubic_derivative_real_ewald.ixx:32-56 with a unit probe
            // charge; the staged charge already carries the Coulomb conversion factor.
            const double ir = rsq64(rrc);
            const double r = rrc * ir;
            const double ir2 = ir * ir;
            const double ir3 = ir2 * ir;
            const double ir5 = ir3 * ir2;
            const double ir7 = ir5 * ir2;
            const double ar2 = A.alpha2 * rrc;
            const double qex = c1.y * expNeg(ar2);                   // q exp(-alpha^2 r^2)
            const double ec = qex * erfcxLds(A.alpha * r, erfcxTab);  // q erfc(alpha r)
            const double etr = A.k2a * r * qex;                     // q 2 alpha / sqrt(pi) r exp(-alpha^2 r^2)
            e = ec * ir;
            f1 = -(etr + ec) * ir3;
            f2 = fma(etr, fma(2.0, ar2, 3.0), 3.0 * ec) * ir5;
            f3 = -fma(etr, fma(ar2, fma(4.0, ar2, 10.0), 15.0), 15.0 * ec) * ir7;
          }
          acc[NG - 1].add(e, f1, f2, f3, u0, u1, u2, u01, u02, u12, u012);
        }
      }
    }
    __syncthreads();
  }

  if (!T.active) return;
  const size_t pidx = 8 * (size_t(T.gi) + size_t(A.n[0]) * (size_t(T.gj) + size_t(A.n[1]) * size_t(T.gk)));
#pragma unroll
  for (int q = 0; q < NG; ++q)
  {
    double v[8];
    v[0] = acc[q].e;
    if (TRIC)
    {
      v[1] = acc[q].d1[0];
      v[2] = acc[q].d1[1];
      v[3] = acc[q].d1[2];
      v[4] = fma(A.G01, acc[q].s1, acc[q].d2[0]);
      v[5] = fma(A.G02, acc[q].s1, acc[q].d2[1]);
      v[6] = fma(A.G12, acc[q].s1, acc[q].d2[2]);
      v[7] = acc[q].d3 + A.G01 * acc[q].b[2] + A.G02 * acc[q].b[1] + A.G12 * acc[q].b[0];
    }
    else
    {
      // src/raspakit/interactions/framework_molecule_grid.cpp:236-243
      v[1] = acc[q].d1[0] * A.diag[0];
      v[2] = acc[q].d1[1] * A.diag[1];
      v[3] = acc[q].d1[2] * A.diag[2];
      v[4] = acc[q].d2[0] * A.diag[0] * A.diag[1];
      v[5] = acc[q].d2[1] * A.diag[0] * A.diag[2];
      v[6] = acc[q].d2[2] * A.diag[1] * A.diag[2];
      v[7] = acc[q].d3 * A.diag[0] * A.diag[1] * A.diag[2];
    }
    // src/raspakit/interpolation_energy_grid.cpp:3849-3874
    v[1] /= A.ncm1[0];
    v[2] /= A.ncm1[1];
    v[3] /= A.ncm1[2];
    v[4] /= A.ncm1[0] * A.ncm1[1];
    v[5] /= A.ncm1[0] * A.ncm1[2];
    v[6] /= A.ncm1[1] * A.ncm1[2];
    v[7] /= A.ncm1[0] * A.ncm1[1] * A.ncm1[2];
    if (v[0] > A.overlap)
    {
      v[0] = 2.0 * A.overlap;
      v[1] = fmin(fmax(v[1], -A.overlap), A.overlap);
      v[2] = fmin(fmax(v[2], -A.overlap), A.overlap);
      v[3] = fmin(fmax(v[3], -A.overlap), A.overlap);
      v[4] = v[5] = v[6] = v[7] = 0.0;
    }
    double2* o = reinterpret_cast<double2*>(A.out[q] + pidx);
    o[0] = make_double2(v[0], v[1]);
    o[1] = make_double2(v[2], v[3]);
    o[2] = make_double2(v[4], v[5]);
    o[3] = make_double2(v[6], v[7]);
  }
}

// ---- triquintic builder: one grid per launch (Lennard-Jones of one probe, or the unit-charge Ewald real-space grid).
// Per pair the radial factors f_0 .. f_6 (f_k = ((1/r) d/dr)^k U) are folded into the moments S(k, e) = sum f_k u^e
// of quintic_terms.hpp; the 27 values are linear combinations of the moments, formed once per point.
//   LJ:    src/raspakit/potentials/interpolation/triquintic_derivative_lj.ixx:33-53
//   Ewald: src/raspakit/potentials/interpolation/triquintic_derivative_real_ewald.ixx:33-80
template <bool LJ, bool TRIC>
__global__ __launch_bounds__(256) void k_grid_build_q(GridArgs A)
{
  constexpr int NLJ = LJ ? 1 : 0;
  constexpr bool COUL = !LJ;
  constexpr int NM = TRIC ? rgq::kMomFull : rgq::kMomDiag;
  __shared__ Img cand[CH];
  __shared__ LJPar par[1][MAXT];
  __shared__ int rowStart[MAXROWS];
  __shared__ int rowPrefix[MAXROWS + 1];
  __shared__ int waveCount[4];
  __shared__ __attribute__((aligned(16))) double erfcxTab[COUL ? kErfcxTabLen : 1];

  const int tid = threadIdx.x;
  if (LJ)
  {
    for (int q = tid; q < MAXT; q += 256) par[0][q] = A.ljpar[q];
  }
  else
  {
    for (int q = tid; q < kErfcxTabLen; q += 256) erfcxTab[q] = erfcxTabEntry(q);
  }
  const TilePoint T = tileSetup<TRIC>(A, rowStart, rowPrefix);
  const int total = rowPrefix[T.nrow];

  double S[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) S[m] = 0.0;

  for (int base = 0; base < total; base += 256)
  {
    const int ncand = stageChunk<NLJ, COUL>(A, T, base, total, rowStart, rowPrefix, par, cand, waveCount);
    const double2* cb = reinterpret_cast<const double2*>(cand);
    for (int c = 0; c < ncand; ++c)
    {
      const double2 c0 = cb[4 * c], c1 = cb[4 * c + 1];  // (x, y), (z, q): an LDS broadcast
      const Img& b = cand[c];
      const double dx = T.px - c0.x, dy = T.py - c0.y, dz = T.pz - c1.x;
      double rr = dx * dx + dy * dy + dz * dz;
      if (rr < A.rc2)
      {
        rr = fmax(rr, 0.1);  // src/raspakit/interactions/framework_molecule_grid.cpp:576
        double f[7];
        if (LJ)
        {
          const LJPar& P = par[0][b.type];
          const double ir2 = rcp64(rr);
          const double x = P.s2 * ir2;
          const double t3 = x * x * x;
          const double a6 = P.e4 * (t3 * t3), a3 = P.e4 * t3;  // 4 eps t^6, 4 eps t^3
          const double ir4 = ir2 * ir2, ir6 = ir4 * ir2, ir8 = ir4 * ir4, ir10 = ir8 * ir2, ir12 = ir6 * ir6;
          f[0] = (a6 - a3) - P.sh;
          f[1] = -6.0 * fma(2.0, a6, -a3) * ir2;
          f[2] = 24.0 * fma(7.0, a6, -2.0 * a3) * ir4;
          f[3] = -96.0 * fma(28.0, a6, -5.0 * a3) * ir6;
          f[4] = 1152.0 * fma(42.0, a6, -5.0 * a3) * ir8;
          f[5] = -80640.0 * fma(12.0, a6, -a3) * ir10;
          f[6] = 645120.0 * fma(33.0, a6, -2.0 * a3) * ir12;
        }
        else
        {
          // unit probe charge; the staged charge carries the Coulomb conversion factor
          const double ir = rsq64(rr);
          const double r = rr * ir;
          const double ir2 = ir * ir;
          const double ar2 = A.alpha2 * rr;
          const double qex = c1.y * expNeg(ar2);                   // q exp(-alpha^2 r^2)
          const double ec = qex * erfcxLds(A.alpha * r, erfcxTab);  // q erfc(alpha r)
          const double etr = A.k2a * r * qex;                     // q 2 alpha / sqrt(pi) r exp(-alpha^2 r^2)
          const double p2 = fma(2.0, ar2, 3.0);
          const double p3 = fma(ar2, fma(4.0, ar2, 10.0), 15.0);
          const double p4 = fma(ar2, fma(ar2, fma(8.0, ar2, 28.0), 70.0), 105.0);
          const double p5 = fma(ar2, fma(ar2, fma(ar2, fma(16.0, ar2, 72.0), 252.0), 630.0), 945.0);
          const double p6 =
              fma(ar2, fma(ar2, fma(ar2, fma(ar2, fma(32.0, ar2, 176.0), 792.0), 2772.0), 6930.0), 10395.0);
          const double ir3 = ir2 * ir, ir5 = ir3 * ir2, ir7 = ir5 * ir2, ir9 = ir7 * ir2, ir11 = ir9 * ir2,
                       ir13 = ir11 * ir2;
          f[0] = ec * ir;
          f[1] = -(etr + ec) * ir3;
          f[2] = fma(etr, p2, 3.0 * ec) * ir5;
          f[3] = -fma(etr, p3, 15.0 * ec) * ir7;
          f[4] = fma(etr, p4, 105.0 * ec) * ir9;
          f[5] = -fma(etr, p5, 945.0 * ec) * ir11;
          f[6] = fma(etr, p6, 10395.0 * ec) * ir13;
        }
        if (TRIC)
          rgq::accumulate<true>(S, f, T.pu0 - b.u0, T.pu1 - b.u1, T.pu2 - b.u2);
        else
          rgq::accumulate<false>(S, f, dx, dy, dz);
      }
    }
    __syncthreads();
  }

  if (!T.active) return;
  double v[27];
  {
    const double G[6] = {A.G00, A.G11, A.G22, A.G01, A.G02, A.G12};
    rgq::combine<TRIC>(S, G, v);
  }
  // rectangular cells: Cartesian derivatives times the cell lengths (framework_molecule_grid.cpp:655-689); then
  // the cell counts (src/raspakit/interpolation_energy_grid.cpp:3888-3924)
#pragma unroll
  for (int o = 1; o < 27; ++o)
  {
    double sc = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      for (int r = 0; r < rgq::kOrders[o][a]; ++r) sc *= TRIC ? 1.0 / A.ncm1[a] : A.diag[a] / A.ncm1[a];
    v[o] *= sc;
  }
  // src/raspakit/interpolation_energy_grid.cpp:3926-3956
  if (v[0] > A.overlap)
  {
    v[0] = 2.0 * A.overlap;
    for (int l = 1; l < 4; ++l) v[l] = fmin(fmax(v[l], -A.overlap), A.overlap);
    for (int l = 4; l < 27; ++l) v[l] = 0.0;
  }
  double* o = A.out[0] + 27 * (size_t(T.gi) + size_t(A.n[0]) * (size_t(T.gj) + size_t(A.n[1]) * size_t(T.gk)));
#pragma unroll
  for (int l = 0; l < 27; ++l) o[l] = v[l];
}

using KernelFn = void (*)(GridArgs);

template <int NLJ, bool COUL>
KernelFn pick2(bool tric)
{
  return tric ? k_grid_build<NLJ, COUL, true> : k_grid_build<NLJ, COUL, false>;
}

KernelFn pick(int nlj, bool coul, bool tric, bool split)
{
  switch (nlj)
  {
    case 0:
      if (coul && split) return tric ? k_grid_build<0, true, true, true> : k_grid_build<0, true, false, true>;
      return coul ? pick2<0, true>(tric) : nullptr;
    case 1: return coul ? pick2<1, true>(tric) : pick2<1, false>(tric);
    case 2: return coul ? pick2<2, true>(tric) : pick2<2, false>(tric);
    case 3: return coul ? pick2<3, true>(tric) : pick2<3, false>(tric);
    case 4: return coul ? pick2<4, true>(tric) : pick2<4, false>(tric);
  }
  return nullptr;
}

// Builds (once per framework and cutoff) the binned periodic images of the unit-cell atoms that can lie within
// rc of a point of the closed unit cell [0,1]^3.
rg_status buildImages(rg_framework* fw, double rc, const std::vector<int>& typeRow, double binTarget)
{
  ImageSet& S = fw->images;
  rg_ctx* ctx = fw->ctx;
  // earlier builds may still read the old images
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (S.d_img) hipFree(S.d_img);
  if (S.d_bin_start) hipFree(S.d_bin_start);
  S.d_img = nullptr;
  S.d_bin_start = nullptr;
  S.rc = -1.0;

  double m[3];
  for (int a = 0; a < 3; ++a)
  {
    m[a] = rc / fw->widths[a] + 1e-9;
    S.lo[a] = -m[a];
    const double span = 1.0 + 2.0 * m[a];
    S.nb[a] = std::max(1, std::min(512, static_cast<int>(span * fw->widths[a] / binTarget)));
    S.bw[a] = span / S.nb[a];
  }
  std::vector<Img> imgs;
  std::vector<int> bin;
  const double* H = fw->cell;
  for (int at = 0; at < fw->n; ++at)
  {
    const double* x = &fw->pos[3 * size_t(at)];
    double f[3];
    for (int a = 0; a < 3; ++a) f[a] = fw->inv[a] * x[0] + fw->inv[3 + a] * x[1] + fw->inv[6 + a] * x[2];
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a)
    {
      lo[a] = static_cast<int>(std::ceil(-m[a] - f[a]));
      hi[a] = static_cast<int>(std::floor(1.0 + m[a] - f[a]));
    }
    for (int n2 = lo[2]; n2 <= hi[2]; ++n2)
      for (int n1 = lo[1]; n1 <= hi[1]; ++n1)
        for (int n0 = lo[0]; n0 <= hi[0]; ++n0)
        {
          Img im{};
          im.x = x[0] + n0 * H[0] + n1 * H[3] + n2 * H[6];
          im.y = x[1] + n0 * H[1] + n1 * H[4] + n2 * H[7];
          im.z = x[2] + n0 * H[2] + n1 * H[5] + n2 * H[8];
          im.q = fw->charge[at];
          im.u0 = H[0] * im.x + H[1] * im.y + H[2] * im.z;
          im.u1 = H[3] * im.x + H[4] * im.y + H[5] * im.z;
          im.u2 = H[6] * im.x + H[7] * im.y + H[8] * im.z;
          im.type = typeRow[fw->type[at]];
          const double fi[3] = {f[0] + n0, f[1] + n1, f[2] + n2};
          int b[3];
          for (int a = 0; a < 3; ++a)
            b[a] = std::max(0, std::min(S.nb[a] - 1, static_cast<int>(std::floor((fi[a] - S.lo[a]) / S.bw[a]))));
          imgs.push_back(im);
          bin.push_back(b[0] + S.nb[0] * (b[1] + S.nb[1] * b[2]));
        }
  }
  const int nbins = S.nb[0] * S.nb[1] * S.nb[2];
  std::vector<int> start(nbins + 1, 0);
  for (int b : bin) start[b + 1]++;
  for (int b = 0; b < nbins; ++b) start[b + 1] += start[b];
  std::vector<Img> sorted(imgs.size());
  {
    std::vector<int> fill(start.begin(), start.end() - 1);
    for (size_t q = 0; q < imgs.size(); ++q) sorted[fill[bin[q]]++] = imgs[q];
  }
  S.n_images = static_cast<int>(sorted.size());
  RG_HIP(ctx, hipMalloc(&S.d_img, std::max<size_t>(1, sorted.size()) * sizeof(Img)));
  RG_HIP(ctx, hipMalloc(&S.d_bin_start, (nbins + 1) * sizeof(int)));
  if (!sorted.empty())
    RG_HIP(ctx, hipMemcpyAsync(S.d_img, sorted.data(), sorted.size() * sizeof(Img), hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(S.d_bin_start, start.data(), (nbins + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  S.rc = rc;
  S.binTarget = binTarget;
  return RG_OK;
}

// Local rows for the framework's distinct pseudo-atom types (typeRow[t] = row of type t, rowType[row] = type).
rg_status frameworkTypeRows(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, std::vector<int>& typeRow,
                            std::vector<int>& rowType, const char* who)
{
  typeRow.assign(ff->n, 0);
  rowType.clear();
  std::vector<int> seen(ff->n, -1);
  for (int t : fw->type)
  {
    if (t < 0 || t >= ff->n) return rg::fail(ctx, RG_ERR_INVALID, "%s: framework type out of range", who);
    if (seen[t] < 0)
    {
      seen[t] = static_cast<int>(rowType.size());
      rowType.push_back(t);
    }
  }
  if (rowType.size() > size_t(MAXT))
    return rg::fail(ctx, RG_ERR_UNSUPPORTED, "%s: more than %d framework pseudo-atom types", who, MAXT);
  for (int t = 0; t < ff->n; ++t) typeRow[t] = seen[t] < 0 ? 0 : seen[t];
  return RG_OK;
}

// Lennard-Jones constants of the probe against each framework type row
LJPar ljPar(const rg_forcefield* ff, int probe, int type)
{
  const size_t k = size_t(probe) * ff->n + type;
  LJPar P{};
  P.e4 = 4.0 * ff->eps[k];
  P.s2 = ff->sigma[k] * ff->sigma[k];
  P.sh = ff->shift[k];
  P.c1 = -6.0 * P.e4;
  P.c2 = 24.0 * P.e4;
  P.c3 = -96.0 * P.e4;
  return P;
}

// g(x) = x erfcx(x) / alpha = x exp(x^2) erfc(x) / alpha for x = alpha r in [x0, x1], as a polynomial of degree
// kSplitDeg in t = u * us + uo (|t| <= 1) of u = 1 / x: the Chebyshev interpolant at kSplitDeg + 1 Chebyshev points
// of u, in monomial form, with FP32 coefficients.  x erfcx(x) tends to 1 / sqrt(pi) with a series in 1 / x^2, so in
// u it needs degree 6 where erfcx itself, linear in x, needs 9 for the same FP32 accuracy.  Returns the largest
// relative error of the FP32-coefficient polynomial (evaluated in FP64) against the FP64 library erfc over 4001
// points of the interval.
double fitErfcxF32(double x0, double x1, double alpha, float* coef, double& us, double& uo)
{
  constexpr int N = kSplitDeg + 1;
  const double u0 = 1.0 / x1, u1 = 1.0 / x0;
  const double mid = 0.5 * (u0 + u1), half = 0.5 * (u1 - u0);
  us = 1.0 / half;
  uo = -mid / half;
  auto g = [alpha](double x) { return x * std::exp(x * x) * std::erfc(x) / alpha; };
  double fval[N], cheb[N];
  for (int k = 0; k < N; ++k) fval[k] = g(1.0 / (mid + half * std::cos(M_PI * (k + 0.5) / N)));
  for (int j = 0; j < N; ++j)
  {
    double sum = 0.0;
    for (int k = 0; k < N; ++k) sum += fval[k] * std::cos(M_PI * j * (k + 0.5) / N);
    cheb[j] = (j == 0 ? 1.0 : 2.0) * sum / N;
  }
  // T_j(t) in monomials by T_{j+1} = 2 t T_j - T_{j-1}
  double mono[N] = {0.0}, tprev[N] = {0.0}, tcur[N] = {0.0}, tnext[N];
  tprev[0] = 1.0;
  tcur[1] = 1.0;
  mono[0] += cheb[0];
  for (int i = 0; i < N; ++i) mono[i] += cheb[1] * tcur[i];
  for (int j = 2; j < N; ++j)
  {
    for (int i = 0; i < N; ++i) tnext[i] = (i > 0 ? 2.0 * tcur[i - 1] : 0.0) - tprev[i];
    for (int i = 0; i < N; ++i)
    {
      mono[i] += cheb[j] * tnext[i];
      tprev[i] = tcur[i];
      tcur[i] = tnext[i];
    }
  }
  for (int i = 0; i < N; ++i) coef[i] = static_cast<float>(mono[i]);
  double worst = 0.0;
  for (int k = 0; k <= 4000; ++k)
  {
    const double x = x0 + (x1 - x0) * k / 4000.0;
    const double t = us / x + uo;
    double p = coef[N - 1];
    for (int i = N - 2; i >= 0; --i) p = p * t + coef[i];
    worst = std::max(worst, std::fabs(p / g(x) - 1.0));
  }
  return worst;
}

rg_status gridBuildMany(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fwc, int n_grids,
                        const rg_grid_type* types, const int* probes, const double* cutoffs, rg_grid* const* grids,
                        bool sync, int precision = RG_GRID_FP64, double splitRadius = 0.0)
{
  if (!ctx || !ff || !fwc || n_grids <= 0 || !types || !probes || !cutoffs || !grids)
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build: invalid arguments");
  rg_framework* fw = const_cast<rg_framework*>(fwc);
  const rg_grid* g0 = grids[0];
  for (int g = 0; g < n_grids; ++g)
  {
    const rg_grid* G = grids[g];
    if (!G || !G->d) return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build: grid %d is null", g);
    if (G->n[0] != g0->n[0] || G->n[1] != g0->n[1] || G->n[2] != g0->n[2])
      return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build: fused grids must have equal dimensions");
    if (G->values != g0->values)
      return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build: fused grids must share the interpolation scheme");
    if (std::memcmp(G->cell, fw->cell, sizeof(G->cell)) != 0)
      return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build: grid %d cell differs from the framework unit cell", g);
    if (types[g] != RG_GRID_LENNARD_JONES && types[g] != RG_GRID_EWALD_REAL)
      return rg::fail(ctx, RG_ERR_UNSUPPORTED, "rg_grid_build: grid type %d not supported", int(types[g]));
    if (types[g] == RG_GRID_LENNARD_JONES && (probes[g] < 0 || probes[g] >= ff->n))
      return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build: probe pseudo-atom %d out of range", probes[g]);
    for (int a = 0; a < 3; ++a)
    {
      const double ncell = std::ceil(2.0 * cutoffs[g] / fw->widths[a]);
      if (ncell * fw->widths[a] < 2.0 * ff->rc_framework)
        return rg::fail(ctx, RG_ERR_UNSUPPORTED,
                        "rg_grid_build: supercell from cutoff %g spans less than twice the VDW cutoff %g",
                        cutoffs[g], ff->rc_framework);
    }
  }
  if (g0->n[0] < 2 || g0->n[1] < 2 || g0->n[2] < 2)
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build: at least 2 points per axis");
  std::vector<int> typeRow, rowType;
  {
    const rg_status st = frameworkTypeRows(ctx, ff, fw, typeRow, rowType, "rg_grid_build");
    if (st != RG_OK) return st;
  }
  RG_ENTER(ctx);
  const int* n = g0->n;
  const double* H = fw->cell;
  const double rc = ff->rc_framework;
  // half-diagonal of the largest tile of points (a tile never spans more points than the grid has)
  double rt = 0.0;
  {
    const int T[3] = {std::min(TI, n[0]), std::min(TJ, n[1]), std::min(TK, n[2])};
    double d[3][3];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) d[a][c] = 0.5 * (T[a] - 1) / double(n[a] - 1) * H[3 * a + c];
    for (int sa = -1; sa <= 1; sa += 2)
      for (int sb = -1; sb <= 1; sb += 2)
      {
        double v[3];
        for (int c = 0; c < 3; ++c) v[c] = d[0][c] + sa * d[1][c] + sb * d[2][c];
        rt = std::max(rt, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
      }
  }
  const double R = rc + rt + 1e-6;
  // bins of at least 2 A, and wide enough that the bin box of a tile spans at most 30 x 30 rows
  const double binTarget = std::max(2.0, 2.0 * R / 28.0);
  if (fw->images.rc != rc || fw->images.binTarget < binTarget)
  {
    rg_status s = buildImages(fw, rc, typeRow, binTarget);
    if (s != RG_OK) return s;
  }

  // split into launches of <= MAXG LJ grids, then the Coulomb grid (at most one) in a launch of its own: the fused
  // LJ + Coulomb kernel needs ~190 VGPRs (two wavefronts per SIMD), the separate kernels ~130 and ~140 (three), and
  // the separate launches measured 15 % faster in total on the configuration-1 grids despite the repeated gather
  std::vector<int> lj, coul;
  for (int g = 0; g < n_grids; ++g) (types[g] == RG_GRID_LENNARD_JONES ? lj : coul).push_back(g);
  if (coul.size() > 1)
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build: at most one Ewald real-space grid per call");

  GridArgs A{};
  std::memcpy(A.cell, fw->cell, sizeof(A.cell));
  for (int a = 0; a < 3; ++a)
  {
    A.n[a] = n[a];
    A.delta[a] = 1.0 / double(n[a] - 1);
    A.ncm1[a] = double(n[a] - 1);
    A.diag[a] = fw->cell[4 * a];
    A.lo[a] = fw->images.lo[a];
    A.bw[a] = fw->images.bw[a];
    A.nb[a] = fw->images.nb[a];
  }
  A.G00 = H[0] * H[0] + H[1] * H[1] + H[2] * H[2];
  A.G11 = H[3] * H[3] + H[4] * H[4] + H[5] * H[5];
  A.G22 = H[6] * H[6] + H[7] * H[7] + H[8] * H[8];
  A.G01 = H[0] * H[3] + H[1] * H[4] + H[2] * H[5];
  A.G02 = H[0] * H[6] + H[1] * H[7] + H[2] * H[8];
  A.G12 = H[3] * H[6] + H[4] * H[7] + H[5] * H[8];
  A.rc2 = rc * rc;
  A.cut2 = R * R;
  for (int a = 0; a < 3; ++a) A.ext[a] = R / fw->widths[a] + 1e-9;
  for (int a = 0; a < 3; ++a)
  {
    const int span = static_cast<int>(std::ceil(2.0 * A.ext[a] / A.bw[a])) + 2;
    if (a > 0 && span * span > MAXROWS)
      return rg::fail(ctx, RG_ERR_UNSUPPORTED, "rg_grid_build: bin box too large for the row table");
  }
  A.ntile[0] = (n[0] + TI - 1) / TI;
  A.ntile[1] = (n[1] + TJ - 1) / TJ;
  A.ntile[2] = (n[2] + TK - 1) / TK;
  A.ntiles = A.ntile[0] * A.ntile[1] * A.ntile[2];
  A.img = static_cast<const Img*>(fw->images.d_img);
  A.bin_start = fw->images.d_bin_start;
  A.alpha = ff->alpha;
  A.alpha2 = ff->alpha * ff->alpha;
  A.k2a = 2.0 * ff->alpha / std::sqrt(M_PI);
  A.coulomb = rg::units().coulomb;
  A.overlap = ff->overlap;
  const bool split = precision == RG_GRID_SPLIT_FP32 && !coul.empty() && splitRadius < rc;
  if (split)
  {
    // images within the split radius of a wavefront keep the FP64 path; the others have alpha r in
    // [alpha rsplit, alpha rc] (with a relative margin of 1e-4 for the FP32 rounding of r)
    double us = 0.0, uo = 0.0;
    const double err = fitErfcxF32(ff->alpha * splitRadius * (1.0 - 1e-4), ff->alpha * rc * (1.0 + 1e-4), ff->alpha,
                                   A.ecx, us, uo);
    if (!(err < 4e-8))
      return rg::fail(ctx, RG_ERR_UNSUPPORTED,
                      "rg_grid_build: the FP32 erfc polynomial reaches only %.2e relative on alpha r in [%g, %g]; "
                      "choose a larger split radius", err, ff->alpha * splitRadius, ff->alpha * rc);
    if (splitRadius * splitRadius < 0.1)
      return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build: a split radius below sqrt(0.1) A lies inside the overlap "
                      "clamp of the pair distance");
    A.rsplit2 = splitRadius * splitRadius;
    A.a2l2e = ff->alpha * ff->alpha / M_LN2;
    // t = u us + uo with u = 1 / (alpha r): the kernel forms t from its FP32 1 / r
    A.xs32 = static_cast<float>(us / ff->alpha);
    A.xo32 = static_cast<float>(uo);
    A.k2a32 = static_cast<float>(A.k2a);
    A.alpha2_32 = static_cast<float>(A.alpha2);
  }

  const bool tric = !fw->rect;
  if (g0->values == 27)
  {
    // triquintic: one launch per grid, the LJ parameter rows of launch l at table offset l * MAXT
    std::vector<LJPar> table(size_t(n_grids) * MAXT, LJPar{});
    for (int g = 0; g < n_grids; ++g)
      if (types[g] == RG_GRID_LENNARD_JONES)
        for (size_t row = 0; row < rowType.size(); ++row)
          table[size_t(g) * MAXT + row] = ljPar(ff, probes[g], rowType[row]);
    const size_t bytes = table.size() * sizeof(LJPar);
    if (ctx->table_bytes < bytes)
    {
      RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
      if (ctx->d_table) hipFree(ctx->d_table);
      ctx->d_table = nullptr;
      ctx->table_bytes = 0;
      RG_HIP(ctx, hipMalloc(&ctx->d_table, bytes));
      ctx->table_bytes = bytes;
    }
    RG_HIP(ctx, hipMemcpyAsync(ctx->d_table, table.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    for (int g = 0; g < n_grids; ++g)
    {
      const bool isLJ = types[g] == RG_GRID_LENNARD_JONES;
      for (int q = 0; q <= MAXG; ++q) A.out[q] = nullptr;
      A.out[0] = grids[g]->d;
      A.ljpar = static_cast<const LJPar*>(ctx->d_table) + size_t(g) * MAXT;
      KernelFn fn = isLJ ? (tric ? k_grid_build_q<true, true> : k_grid_build_q<true, false>)
                         : (tric ? k_grid_build_q<false, true> : k_grid_build_q<false, false>);
      hipLaunchKernelGGL(fn, dim3(A.ntiles), dim3(256), 0, ctx->stream, A);
      RG_HIP(ctx, hipGetLastError());
    }
    if (sync) RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RG_OK;
  }
  // parameter tables of all launches, copied once into the context's persistent table buffer
  struct Launch
  {
    int nlj;
    bool coul;
    double* out[MAXG + 1];
  };
  std::vector<Launch> launches;
  std::vector<LJPar> table;
  {
    size_t li = 0;
    bool coulDone = coul.empty();
    while (li < lj.size() || !coulDone)
    {
      Launch L{};
      L.nlj = static_cast<int>(std::min<size_t>(MAXG, lj.size() - li));
      L.coul = L.nlj == 0 && !coulDone;
      const size_t t0 = table.size();
      table.resize(t0 + MAXG * MAXT, LJPar{});
      for (int q = 0; q < L.nlj; ++q)
      {
        const int g = lj[li + q];
        for (size_t row = 0; row < rowType.size(); ++row) table[t0 + q * MAXT + row] = ljPar(ff, probes[g], rowType[row]);
        L.out[q] = grids[g]->d;
      }
      if (L.coul) L.out[L.nlj] = grids[coul[0]]->d;
      launches.push_back(L);
      li += L.nlj;
      coulDone = coulDone || L.coul;
    }
  }
  const size_t bytes = table.size() * sizeof(LJPar);
  if (ctx->table_bytes < bytes)
  {
    // waits for earlier launches that still read the old table before it is released
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->d_table) hipFree(ctx->d_table);
    ctx->d_table = nullptr;
    ctx->table_bytes = 0;
    RG_HIP(ctx, hipMalloc(&ctx->d_table, bytes));
    ctx->table_bytes = bytes;
  }
  RG_HIP(ctx, hipMemcpyAsync(ctx->d_table, table.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
  for (size_t l = 0; l < launches.size(); ++l)
  {
    const Launch& L = launches[l];
    for (int q = 0; q <= MAXG; ++q) A.out[q] = L.out[q];
    A.ljpar = static_cast<const LJPar*>(ctx->d_table) + l * MAXG * MAXT;
    KernelFn fn = pick(L.nlj, L.coul, tric, split);
    if (!fn) return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build: nothing to build");
    hipLaunchKernelGGL(fn, dim3(A.ntiles), dim3(256), 0, ctx->stream, A);
    RG_HIP(ctx, hipGetLastError());
  }
  if (sync) RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}


// ---- grid check: the test of System::createFrameworkInterpolationGrids on random points
// (src/raspakit/system_interpolation_grids.cpp:240-840).  Per point the analytic Lennard-Jones and unit-charge Ewald
// real-space energies and Cartesian gradients (sum f1 dr over the in-cutoff images, the same pair terms as the
// builder) are compared with the grids' tricubic interpolation.
struct CheckArgs
{
  const Img* img;
  int n_images;
  const LJPar* par;  // [MAXT] rows of the probe
  const double* pos;  // n * 3, Cartesian, inside the unit cell
  int n;
  int coul;
  double rc2, alpha, alpha2, k2a, coulomb;
  double* out;  // n * 20: E_lj, dE_lj/dr (3), d2E_lj/dr2 (xx xy xz yy yz zz), then the same for a unit probe charge
};

__global__ __launch_bounds__(256) void k_grid_check(CheckArgs A)
{
  __shared__ __attribute__((aligned(16))) double erfcxTab[kErfcxTabLen];
  for (int q = threadIdx.x; q < kErfcxTabLen; q += 256) erfcxTab[q] = erfcxTabEntry(q);
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool live = i < A.n;
  const double px = live ? A.pos[3 * i] : 0.0, py = live ? A.pos[3 * i + 1] : 0.0, pz = live ? A.pos[3 * i + 2] : 0.0;
  // per grid type: energy, gradient, Hessian (xx xy xz yy yz zz); Cartesian second derivatives f2 dr_a dr_b + f1 d_ab
  double v[2][10] = {};
  auto add = [&](double* o, double e, double f1, double f2, double dx, double dy, double dz)
  {
    o[0] += e;
    o[1] = fma(f1, dx, o[1]);
    o[2] = fma(f1, dy, o[2]);
    o[3] = fma(f1, dz, o[3]);
    o[4] += fma(f2, dx * dx, f1);
    o[5] = fma(f2, dx * dy, o[5]);
    o[6] = fma(f2, dx * dz, o[6]);
    o[7] += fma(f2, dy * dy, f1);
    o[8] = fma(f2, dy * dz, o[8]);
    o[9] += fma(f2, dz * dz, f1);
  };
  for (int j = 0; j < A.n_images; ++j)  // uniform over the wavefront: every lane reads the same image
  {
    const Img& b = A.img[j];
    const double dx = px - b.x, dy = py - b.y, dz = pz - b.z;
    double rr = dx * dx + dy * dy + dz * dz;
    if (live && rr < A.rc2)
    {
      rr = fmax(rr, 0.1);  // src/raspakit/interactions/framework_molecule_grid.cpp:200
      const LJPar& P = A.par[b.type];
      if (P.e4 != 0.0)
      {
        // src/raspakit/potentials/interpolation/tricubic_derivative_lj.ixx:33-46 (energy, f1, f2)
        const double ir2 = rcp64(rr);
        const double x = P.s2 * ir2;
        const double t3 = x * x * x;
        const double t6 = t3 * t3;
        const double f1 = P.c1 * fma(2.0, t6, -t3) * ir2;
        const double f2 = P.c2 * fma(7.0, t6, -2.0 * t3) * ir2 * ir2;
        add(v[0], fma(P.e4, t6 - t3, -P.sh), f1, f2, dx, dy, dz);
      }
      if (A.coul && b.q != 0.0)
      {
        // src/raspakit/potentials/interpolation/tricubic_derivative_real_ewald.ixx:32-56 (energy, f1, f2)
        const double ir = rsq64(rr);
        const double r = rr * ir;
        const double q = b.q * A.coulomb;
        const double ar2 = A.alpha2 * rr;
        const double qex = q * expNeg(ar2);
        const double ec = qex * erfcxLds(A.alpha * r, erfcxTab);
        const double etr = A.k2a * r * qex;
        const double ir3 = ir * ir * ir;
        const double f1 = -(etr + ec) * ir3;
        const double f2 = fma(etr, fma(2.0, ar2, 3.0), 3.0 * ec) * ir3 * ir * ir;
        add(v[1], ec * ir, f1, f2, dx, dy, dz);
      }
    }
  }
  if (!live) return;
  double* o = A.out + 20 * size_t(i);
  for (int k = 0; k < 10; ++k)
  {
    o[k] = v[0][k];
    o[10 + k] = v[1][k];
  }
}

// Boltzmann-weighted comparison of 'full' (analytic) and 'table' (interpolated) values
struct CheckSums
{
  double w = 0, table = 0, full = 0, diff2 = 0, full2 = 0;
  double gt[3] = {0, 0, 0}, gf[3] = {0, 0, 0}, gd2[3] = {0, 0, 0}, gf2[3] = {0, 0, 0};
  double ht[6] = {}, hf[6] = {}, hd2[6] = {}, hf2[6] = {};
  bool hess = false;
  // Hessians as 6 components xx xy xz yy yz zz (system_interpolation_grids.cpp:404-441)
  void addHessian(double wt, const double* htab, const double* hfull)
  {
    hess = true;
    for (int a = 0; a < 6; ++a)
    {
      ht[a] += htab[a] * wt;
      hf[a] += hfull[a] * wt;
      hd2[a] += (hfull[a] - htab[a]) * (hfull[a] - htab[a]) * wt;
      hf2[a] += hfull[a] * hfull[a] * wt;
    }
  }
  void add(double wt, double et, double ef, const double* gtab, const double* gfull)
  {
    w += wt;
    table += et * wt;
    full += ef * wt;
    diff2 += (ef - et) * (ef - et) * wt;
    full2 += ef * ef * wt;
    for (int a = 0; a < 3; ++a)
    {
      gt[a] += gtab[a] * wt;
      gf[a] += gfull[a] * wt;
      gd2[a] += (gfull[a] - gtab[a]) * (gfull[a] - gtab[a]) * wt;
      gf2[a] += gfull[a] * gfull[a] * wt;
    }
  }
  void store(rg_grid_check_stats* s) const
  {
    s->weight = w;
    s->energy_table = table / w;
    s->energy_full = full / w;
    s->energy_relative_error = std::sqrt(diff2 / full2);
    for (int a = 0; a < 3; ++a)
    {
      s->gradient_table[a] = gt[a] / w;
      s->gradient_full[a] = gf[a] / w;
      s->gradient_relative_error[a] = std::sqrt(gd2[a] / gf2[a]);
    }
    if (hess)
      for (int a = 0; a < 6; ++a)
      {
        s->hessian_table[a] = ht[a] / w;
        s->hessian_full[a] = hf[a] / w;
        s->hessian_relative_error[a] = std::sqrt(hd2[a] / hf2[a]);
      }
  }
};

rg_status gridCheck(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fwc, const rg_grid* lj, int probe,
                    const rg_grid* coul, double probeCharge, double temperature, int n, const double* frac,
                    unsigned long long seed, rg_grid_check_t* out)
{
  const char* who = "rg_grid_check";
  if (!ctx || !ff || !fwc || !lj || !out || n <= 0 || temperature <= 0.0)
    return rg::fail(ctx, RG_ERR_INVALID, "%s: invalid arguments", who);
  if (probe < 0 || probe >= ff->n) return rg::fail(ctx, RG_ERR_INVALID, "%s: probe pseudo-atom out of range", who);
  rg_framework* fw = const_cast<rg_framework*>(fwc);
  for (const rg_grid* g : {lj, coul})
  {
    if (g && std::memcmp(g->cell, fw->cell, sizeof(g->cell)) != 0)
      return rg::fail(ctx, RG_ERR_INVALID, "%s: grid cell differs from the framework unit cell", who);
  }
  if (coul && coul->values != lj->values)
    return rg::fail(ctx, RG_ERR_INVALID, "%s: the LJ and Coulomb grids must share the interpolation scheme", who);
  const bool quintic = lj->values == 27;
  std::vector<int> typeRow, rowType;
  {
    const rg_status st = frameworkTypeRows(ctx, ff, fw, typeRow, rowType, who);
    if (st != RG_OK) return st;
  }
  RG_ENTER(ctx);
  const double rc = ff->rc_framework;
  if (fw->images.rc != rc)
  {
    const rg_status st = buildImages(fw, rc, typeRow, 2.0);
    if (st != RG_OK) return st;
  }
  // test points: uniform in the unit cell (the caller's fractional coordinates, or a Philox4x32-10 stream)
  std::vector<double> pos(3 * size_t(n));
  for (int i = 0; i < n; ++i)
  {
    double s[3];
    if (frac)
      for (int a = 0; a < 3; ++a) s[a] = frac[3 * size_t(i) + a];
    else
    {
      double u3;
      rgd::uniform2(seed, uint32_t(i), 0x67726964u, 0u, 0u, s[0], s[1]);
      rgd::uniform2(seed, uint32_t(i), 0x67726964u, 1u, 0u, s[2], u3);
    }
    for (int c = 0; c < 3; ++c)
      pos[3 * size_t(i) + c] = fw->cell[c] * s[0] + fw->cell[3 + c] * s[1] + fw->cell[6 + c] * s[2];
  }
  std::vector<double> eLJ(n), gLJ(3 * size_t(n)), hLJ, eQ, gQ, hQ, full(20 * size_t(n));
  auto interp = [&](const rg_grid* g, std::vector<double>& e, std::vector<double>& gr, std::vector<double>& h)
  {
    e.resize(n);
    gr.resize(3 * size_t(n));
    if (!quintic) return rg_grid_interpolate(ctx, g, n, pos.data(), e.data(), gr.data());
    h.resize(9 * size_t(n));
    return rg_grid_interpolate_hessian(ctx, g, n, pos.data(), e.data(), gr.data(), h.data());
  };
  {
    const rg_status st = interp(lj, eLJ, gLJ, hLJ);
    if (st != RG_OK) return st;
  }
  if (coul)
  {
    const rg_status st = interp(coul, eQ, gQ, hQ);
    if (st != RG_OK) return st;
  }
  std::vector<LJPar> par(MAXT);
  for (size_t row = 0; row < rowType.size(); ++row) par[row] = ljPar(ff, probe, rowType[row]);
  void *dPar = nullptr, *dPos = nullptr, *dOut = nullptr;
  auto release = [&]
  {
    hipStreamSynchronize(ctx->stream);
    if (dPar) hipFree(dPar);
    if (dPos) hipFree(dPos);
    if (dOut) hipFree(dOut);
  };
  auto run = [&]() -> rg_status
  {
    RG_HIP(ctx, hipMalloc(&dPar, MAXT * sizeof(LJPar)));
    RG_HIP(ctx, hipMalloc(&dPos, pos.size() * sizeof(double)));
    RG_HIP(ctx, hipMalloc(&dOut, full.size() * sizeof(double)));
    RG_HIP(ctx, hipMemcpyAsync(dPar, par.data(), MAXT * sizeof(LJPar), hipMemcpyHostToDevice, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(dPos, pos.data(), pos.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    CheckArgs A{};
    A.img = static_cast<const Img*>(fw->images.d_img);
    A.n_images = fw->images.n_images;
    A.par = static_cast<const LJPar*>(dPar);
    A.pos = static_cast<const double*>(dPos);
    A.n = n;
    A.coul = coul ? 1 : 0;
    A.rc2 = rc * rc;
    A.alpha = ff->alpha;
    A.alpha2 = ff->alpha * ff->alpha;
    A.k2a = 2.0 * ff->alpha / std::sqrt(M_PI);
    A.coulomb = rg::units().coulomb;
    A.out = static_cast<double*>(dOut);
    hipLaunchKernelGGL(k_grid_check, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, A);
    RG_HIP(ctx, hipGetLastError());
    RG_HIP(ctx, hipMemcpyAsync(full.data(), dOut, full.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RG_OK;
  };
  const rg_status st = run();
  release();
  if (st != RG_OK) return st;

  // Statistics as the reference accumulates them (energies and gradients in K and K/A): every point is weighted
  // by exp(-beta E_LJ) with beta = 1 / (k_B T) in inverse internal energy units and E_LJ, the analytic
  // Lennard-Jones energy, in Kelvin (system_interpolation_grids.cpp:365, :545); the Ewald comparison, scaled by the
  // probe's charge, takes the points whose analytic E_LJ (K) is below the overlap criterion (:420).
  const rg::Units& U = rg::units();
  const double beta = 1.0 / (U.kB * temperature);
  const double K = U.energyToKelvin;
  CheckSums sLJ, sQ;
  long long nQ = 0;
  // the 6 independent components of an interpolated (row-major 3x3) Hessian, times a scale
  auto six = [](const double* h9, double sc, double* h6)
  {
    h6[0] = h9[0] * sc;
    h6[1] = h9[1] * sc;
    h6[2] = h9[2] * sc;
    h6[3] = h9[4] * sc;
    h6[4] = h9[5] * sc;
    h6[5] = h9[8] * sc;
  };
  for (int i = 0; i < n; ++i)
  {
    const double* f = &full[20 * size_t(i)];
    const double ef = f[0] * K, et = eLJ[i] * K;
    const double gf[3] = {f[1] * K, f[2] * K, f[3] * K};
    const double gt[3] = {gLJ[3 * size_t(i)] * K, gLJ[3 * size_t(i) + 1] * K, gLJ[3 * size_t(i) + 2] * K};
    const double w = std::exp(-beta * ef);
    sLJ.add(w, et, ef, gt, gf);
    if (quintic)
    {
      double ht[6], hf[6];
      six(&hLJ[9 * size_t(i)], K, ht);
      for (int a = 0; a < 6; ++a) hf[a] = f[4 + a] * K;
      sLJ.addHessian(w, ht, hf);
    }
    if (coul && ef < ff->overlap)
    {
      const double qK = probeCharge * K;
      const double qf[3] = {f[11] * qK, f[12] * qK, f[13] * qK};
      const double qt[3] = {gQ[3 * size_t(i)] * qK, gQ[3 * size_t(i) + 1] * qK, gQ[3 * size_t(i) + 2] * qK};
      sQ.add(w, eQ[i] * qK, f[10] * qK, qt, qf);
      if (quintic)
      {
        double ht[6], hf[6];
        six(&hQ[9 * size_t(i)], qK, ht);
        for (int a = 0; a < 6; ++a) hf[a] = f[14 + a] * qK;
        sQ.addHessian(w, ht, hf);
      }
      ++nQ;
    }
  }
  *out = rg_grid_check_t{};
  out->n_points = n;
  sLJ.store(&out->lennard_jones);
  out->n_coulomb_points = nQ;
  if (coul && nQ > 0) sQ.store(&out->coulomb);
  return RG_OK;
}
}  // namespace

extern "C"
{

rg_status rg_grid_create(rg_ctx* ctx, const double* cell9, int is_rectangular, int nx, int ny, int nz, rg_grid** out)
{
  return rg_grid_create_scheme(ctx, cell9, is_rectangular, nx, ny, nz, RG_SCHEME_TRICUBIC, out);
}

rg_status rg_grid_create_scheme(rg_ctx* ctx, const double* cell9, int is_rectangular, int nx, int ny, int nz,
                                rg_interpolation_scheme scheme, rg_grid** out)
{
  if (!ctx || !cell9 || !out || nx < 2 || ny < 2 || nz < 2)
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_create: invalid arguments");
  if (scheme != RG_SCHEME_TRICUBIC && scheme != RG_SCHEME_TRIQUINTIC)
    return rg::fail(ctx, RG_ERR_UNSUPPORTED, "rg_grid_create: interpolation scheme %d not supported", int(scheme));
  rg_grid* g = new rg_grid();
  g->ctx = ctx;
  g->device = ctx->device;
  std::memcpy(g->cell, cell9, sizeof(g->cell));
  rg::invert3(g->cell, g->inv);
  g->rect = is_rectangular;
  g->n[0] = nx;
  g->n[1] = ny;
  g->n[2] = nz;
  g->values = int(scheme);
  g->count = size_t(g->values) * size_t(nx) * size_t(ny) * size_t(nz);
  hipSetDevice(ctx->device);
  hipError_t e = hipMalloc(&g->d, g->count * sizeof(double));
  if (e != hipSuccess)
  {
    delete g;
    return rg::hipfail(ctx, e, "hipMalloc(grid)");
  }
  *out = g;
  return RG_OK;
}

rg_status rg_grid_destroy(rg_grid* g)
{
  if (!g) return RG_ERR_INVALID;
  rg_status st = rgReleaseStatus(hipSetDevice(g->device), RG_OK);
  if (g->d) st = rgReleaseStatus(hipFree(g->d), st);
  delete g;
  return st;
}

void* rg_grid_device_data(rg_grid* g) { return g ? g->d : nullptr; }

rg_status rg_grid_build(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, rg_grid_type type,
                        int probe_pseudo_atom, double cutoff_supercell, rg_grid* grid)
{
  rg_grid* gs[1] = {grid};
  return gridBuildMany(ctx, ff, fw, 1, &type, &probe_pseudo_atom, &cutoff_supercell, gs, true);
}

rg_status rg_grid_build_many(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, int n_grids,
                             const rg_grid_type* types, const int* probes, const double* cutoffs,
                             rg_grid* const* grids)
{
  return gridBuildMany(ctx, ff, fw, n_grids, types, probes, cutoffs, grids, true);
}

rg_status rg_grid_build_many_async(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, int n_grids,
                                   const rg_grid_type* types, const int* probes, const double* cutoffs,
                                   rg_grid* const* grids)
{
  return gridBuildMany(ctx, ff, fw, n_grids, types, probes, cutoffs, grids, false);
}

rg_status rg_grid_build_many_precision(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, int n_grids,
                                       const rg_grid_type* types, const int* probes, const double* cutoffs,
                                       rg_grid* const* grids, rg_grid_precision precision, double split_radius,
                                       int async)
{
  if (precision != RG_GRID_FP64 && precision != RG_GRID_SPLIT_FP32)
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build_many_precision: unknown precision %d", int(precision));
  if (precision == RG_GRID_SPLIT_FP32 && !(split_radius > 0.0))
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_build_many_precision: the split radius must be positive");
  return gridBuildMany(ctx, ff, fw, n_grids, types, probes, cutoffs, grids, async == 0, int(precision), split_radius);
}

rg_status rg_grid_check(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, const rg_grid* lj_grid, int probe,
                        const rg_grid* coulomb_grid, double probe_charge, double temperature, int n_points,
                        const double* frac, unsigned long long seed, rg_grid_check_t* out)
{
  return gridCheck(ctx, ff, fw, lj_grid, probe, coulomb_grid, probe_charge, temperature, n_points, frac, seed, out);
}

rg_status rg_grid_download(rg_ctx* ctx, const rg_grid* grid, size_t offset, size_t count, double* host)
{
  if (!ctx || !grid || !host || offset + count > grid->count)
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_download: range out of bounds");
  RG_ENTER(ctx);
  RG_HIP(ctx, hipMemcpyAsync(host, grid->d + offset, count * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

rg_status rg_grid_upload(rg_ctx* ctx, rg_grid* grid, size_t offset, size_t count, const double* host)
{
  if (!ctx || !grid || !host || offset + count > grid->count)
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_upload: range out of bounds");
  RG_ENTER(ctx);
  RG_HIP(ctx, hipMemcpyAsync(grid->d + offset, host, count * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

}  // extern "C"

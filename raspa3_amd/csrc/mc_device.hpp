// Device helpers shared by the trial-energy, Ewald and Monte Carlo kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rgd
{

// ---------------------------------------------------------------------------------------------------
// Simulation box (cell columns = lattice vectors) and the reference's minimum-image rounding
// (src/raspakit/simulationbox.ixx:131-161: truncation of s +- 0.5).
// ---------------------------------------------------------------------------------------------------
struct Box
{
  double cell[9];
  double inv[9];
  int rect;
  int pad;
  double volume;
};

__device__ __forceinline__ double roundHalfAway(double s) { return double((long long)(s + ((s >= 0.0) ? 0.5 : -0.5))); }

__device__ __forceinline__ void pbc(const Box& b, double& dx, double& dy, double& dz)
{
  if (b.rect)
  {
    dx -= roundHalfAway(dx * b.inv[0]) * b.cell[0];
    dy -= roundHalfAway(dy * b.inv[4]) * b.cell[4];
    dz -= roundHalfAway(dz * b.inv[8]) * b.cell[8];
  }
  else
  {
    double s0 = b.inv[0] * dx + b.inv[3] * dy + b.inv[6] * dz;
    double s1 = b.inv[1] * dx + b.inv[4] * dy + b.inv[7] * dz;
    double s2 = b.inv[2] * dx + b.inv[5] * dy + b.inv[8] * dz;
    s0 -= roundHalfAway(s0);
    s1 -= roundHalfAway(s1);
    s2 -= roundHalfAway(s2);
    dx = b.cell[0] * s0 + b.cell[3] * s1 + b.cell[6] * s2;
    dy = b.cell[1] * s0 + b.cell[4] * s1 + b.cell[7] * s2;
    dz = b.cell[2] * s0 + b.cell[5] * s1 + b.cell[8] * s2;
  }
}

// ---------------------------------------------------------------------------------------------------
// Pair energies at full derivative order 0.
// LJ: src/raspakit/potentials/pair_vdw.ixx:74-111 (soft-core form, scaling = lambda_A lambda_B).
// Ewald real space: src/raspakit/potentials/pair_coulomb.ixx:39-76 with the 0.01 A charge offset
// (src/raspakit/potentials/coulomb_real_space.ixx:25).
// ---------------------------------------------------------------------------------------------------
struct Pair
{
  double e4, s2, shift, lj;  // 4 eps, sigma^2, shift (internal units), 1 = LJ / 0 = None
};

__device__ __forceinline__ double vdw0(const Pair& p, double scaling, double rr)
{
  if (p.lj == 0.0) return 0.0;
  const double temp = rr / p.s2;
  const double temp3 = temp * temp * temp;
  const double inv_scaling = 1.0 - scaling;
  const double rri3 = 1.0 / (temp3 + 0.5 * inv_scaling * inv_scaling);
  const double term = p.e4 * (rri3 * (rri3 - 1.0)) - p.shift;
  return scaling * term;
}

__device__ __forceinline__ double coul0(double alpha, double scaling, double r, double Cqq)
{
  const double s = r + 0.01 * (1.0 - scaling);
  return scaling * Cqq * erfc(alpha * s) / s;
}

// ---------------------------------------------------------------------------------------------------
// Tricubic (and triquintic) Hermite evaluation of one cell (see interpolate.hip for the derivation).
// ---------------------------------------------------------------------------------------------------
struct GridView
{
  const double* data;
  double inv[9];
  int n[3];
  double ncd[3];
  int values;  // 8 (tricubic) or 27 (triquintic)
};

__device__ __forceinline__ void hermite(double t, double w[2][2])
{
  const double t2 = t * t, t3 = t2 * t;
  w[0][0] = 2.0 * t3 - 3.0 * t2 + 1.0;
  w[0][1] = t3 - 2.0 * t2 + t;
  w[1][0] = -2.0 * t3 + 3.0 * t2;
  w[1][1] = t3 - t2;
}

// Triquintic energy of one cell: tensor product of 1-D quintic Hermite bases (value, first and second derivative at
// both ends); value l of a record carries the derivative orders (ox, oy, oz) of the reference's layout
// (src/raspakit/interpolation_energy_grid.cpp:3886-3924).
__device__ __forceinline__ void hermite5(double t, double w[2][3])
{
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  w[0][0] = 1.0 - 10.0 * t3 + 15.0 * t4 - 6.0 * t5;
  w[0][1] = t - 6.0 * t3 + 8.0 * t4 - 3.0 * t5;
  w[0][2] = 0.5 * t2 - 1.5 * t3 + 1.5 * t4 - 0.5 * t5;
  w[1][0] = 10.0 * t3 - 15.0 * t4 + 6.0 * t5;
  w[1][1] = -4.0 * t3 + 7.0 * t4 - 3.0 * t5;
  w[1][2] = 0.5 * t3 - t4 + 0.5 * t5;
}

__device__ __forceinline__ double interpolateQuintic(const GridView& g, const size_t c0[3], const double t[3])
{
  constexpr int kOrd[27][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {1, 1, 0}, {1, 0, 1},
                               {0, 2, 0}, {0, 1, 1}, {0, 0, 2}, {2, 1, 0}, {2, 0, 1}, {1, 2, 0}, {1, 1, 1},
                               {0, 2, 1}, {1, 0, 2}, {0, 1, 2}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {2, 1, 1},
                               {1, 2, 1}, {1, 1, 2}, {2, 2, 1}, {2, 1, 2}, {1, 2, 2}, {2, 2, 2}};
  double wx[2][3], wy[2][3], wz[2][3];
  hermite5(t[0], wx);
  hermite5(t[1], wy);
  hermite5(t[2], wz);
  double e = 0.0;
  for (int c = 0; c < 8; ++c)
  {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    const double* F = g.data + 27 * ((c0[0] + cx) + size_t(g.n[0]) * ((c0[1] + cy) + size_t(g.n[1]) * (c0[2] + cz)));
    double X[3], Y[3], Z[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      X[d] = cx ? wx[1][d] : wx[0][d];
      Y[d] = cy ? wy[1][d] : wy[0][d];
      Z[d] = cz ? wz[1][d] : wz[0][d];
    }
#pragma unroll
    for (int l = 0; l < 27; ++l) e = fma(F[l], X[kOrd[l][0]] * Y[kOrd[l][1]] * Z[kOrd[l][2]], e);
  }
  return e;
}

__device__ __forceinline__ double interpolate(const GridView& g, double px, double py, double pz)
{
  double t[3];
  size_t c0[3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
  {
    double v = g.inv[a] * px + g.inv[3 + a] * py + g.inv[6 + a] * pz;
    v -= rint(v);
    if (v < 0.0) v += 1.0;
    if (v > 1.0) v -= 1.0;
    const size_t nc = size_t(g.n[a] - 1);
    size_t c = static_cast<size_t>(v * g.ncd[a]);
    c = c < nc - 1 ? c : nc - 1;
    c0[a] = c;
    t[a] = v * g.ncd[a] - double(c);
  }
  if (g.values == 27) return interpolateQuintic(g, c0, t);
  double wx[2][2], wy[2][2], wz[2][2];
  hermite(t[0], wx);
  hermite(t[1], wy);
  hermite(t[2], wz);
  double e = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c)
  {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    const size_t idx = 8 * ((c0[0] + cx) + size_t(g.n[0]) * ((c0[1] + cy) + size_t(g.n[1]) * (c0[2] + cz)));
    const double2* p2 = reinterpret_cast<const double2*>(g.data + idx);
    const double2 r0 = p2[0], r1 = p2[1], r2 = p2[2], r3 = p2[3];
    const double* x = wx[cx];
    const double* y = wy[cy];
    const double* z = wz[cz];
    const double a00 = fma(r0.x, z[0], r1.y * z[1]);
    const double a01 = fma(r1.x, z[0], r3.x * z[1]);
    const double a10 = fma(r0.y, z[0], r2.y * z[1]);
    const double a11 = fma(r2.x, z[0], r3.y * z[1]);
    e += fma(x[0], fma(y[0], a00, y[1] * a01), x[1] * fma(y[0], a10, y[1] * a11));
  }
  return e;
}

// ---------------------------------------------------------------------------------------------------
// Philox4x32-10 counter-based generator (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11).  The same function runs in the oracle so that trial positions can be
// replayed on the CPU.
// ---------------------------------------------------------------------------------------------------
struct U4
{
  uint32_t x, y, z, w;
};

__host__ __device__ inline U4 philox4x32(U4 c, uint32_t k0, uint32_t k1)
{
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r)
  {
    const uint64_t p0 = uint64_t(M0) * c.x;
    const uint64_t p1 = uint64_t(M1) * c.z;
    const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Two uniforms in [0, 1) with 53 random bits each, from stream (seed) position (a, b, c, d).
__host__ __device__ inline void uniform2(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d, double& u0,
                                         double& u1)
{
  const U4 r = philox4x32(U4{a, b, c, d}, uint32_t(seed), uint32_t(seed >> 32));
  const uint64_t v0 = (uint64_t(r.x) << 32 | r.y) >> 11;
  const uint64_t v1 = (uint64_t(r.z) << 32 | r.w) >> 11;
  u0 = double(v0) * (1.0 / 9007199254740992.0);
  u1 = double(v1) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------------------------------
// Wavefront (64-lane) reductions.
// ---------------------------------------------------------------------------------------------------
__device__ __forceinline__ double waveSum(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double waveMax(double v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

}  // namespace rgd

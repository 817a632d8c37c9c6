// Batched tricubic and triquintic interpolation on gfx950.
//
// Replaces InterpolationEnergyGrid::interpolate / interpolateGradient / interpolateHessian, tricubic and triquintic
// branches (src/raspakit/interpolation_energy_grid.cpp:3980-4180, :4189-4440, :4443-4760).  The reference gathers the
// 8 corner records X (64 or 216 values), forms the polynomial coefficients a = A X with a fixed matrix
// (src/raspakit/interpolation_energy_grid.cpp:193-358 tricubic, :2116-... triquintic) and sums a_ijk x^i y^j z^k.
// Those polynomials are the unique tricubic (triquintic) matching the stored derivatives at the 8 corners, i.e. the
// tensor products of 1-D cubic (quintic) Hermite interpolants; these kernels evaluate them in that form, so the
// results equal the reference's up to rounding.  The Hessian is the symmetric matrix of second derivatives of the
// interpolant (the reference's triquintic branch fills its (z, y) entry with the (x, z) derivative,
// interpolation_energy_grid.cpp:4748); its tricubic branch returns a zero Hessian.
//
// One lane per position.  Tricubic corner records are 64-byte contiguous chunks (component-fastest layout) read
// as four 16-byte loads; triquintic records are 216 bytes read as 27 doubles.
#include <cmath>
#include <cstring>

#include "quintic_terms.hpp"
#include "rg_internal.hpp"

namespace rg
{

struct InterpArgs
{
  double inv[9];
  int n[3];
  double ncd[3];  // number of cells (n - 1) as double
  const double* grid;
  const double* pos;  // n*3
  int count;
  double* energy;
  double* grad;  // may be null
  double* hess;  // may be null (row-major 3x3 per position)
};

// Hermite basis on [0,1]: w[c][d] = weight of corner c for derivative order d; dw, d2w its first and second
// derivatives in t
__device__ __forceinline__ void hermite(double t, double w[2][2], double dw[2][2], double d2w[2][2])
{
  const double t2 = t * t, t3 = t2 * t;
  w[0][0] = 2.0 * t3 - 3.0 * t2 + 1.0;
  w[0][1] = t3 - 2.0 * t2 + t;
  w[1][0] = -2.0 * t3 + 3.0 * t2;
  w[1][1] = t3 - t2;
  dw[0][0] = 6.0 * t2 - 6.0 * t;
  dw[0][1] = 3.0 * t2 - 4.0 * t + 1.0;
  dw[1][0] = -6.0 * t2 + 6.0 * t;
  dw[1][1] = 3.0 * t2 - 2.0 * t;
  d2w[0][0] = 12.0 * t - 6.0;
  d2w[0][1] = 6.0 * t - 4.0;
  d2w[1][0] = -12.0 * t + 6.0;
  d2w[1][1] = 6.0 * t - 2.0;
}

// Quintic Hermite basis on [0,1] (value, first and second derivative at each end) and its t-derivatives
__device__ __forceinline__ void hermite5(double t, double w[2][3], double dw[2][3], double d2w[2][3])
{
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  w[0][0] = 1.0 - 10.0 * t3 + 15.0 * t4 - 6.0 * t5;
  w[0][1] = t - 6.0 * t3 + 8.0 * t4 - 3.0 * t5;
  w[0][2] = 0.5 * t2 - 1.5 * t3 + 1.5 * t4 - 0.5 * t5;
  w[1][0] = 10.0 * t3 - 15.0 * t4 + 6.0 * t5;
  w[1][1] = -4.0 * t3 + 7.0 * t4 - 3.0 * t5;
  w[1][2] = 0.5 * t3 - t4 + 0.5 * t5;
  dw[0][0] = -30.0 * t2 + 60.0 * t3 - 30.0 * t4;
  dw[0][1] = 1.0 - 18.0 * t2 + 32.0 * t3 - 15.0 * t4;
  dw[0][2] = t - 4.5 * t2 + 6.0 * t3 - 2.5 * t4;
  dw[1][0] = 30.0 * t2 - 60.0 * t3 + 30.0 * t4;
  dw[1][1] = -12.0 * t2 + 28.0 * t3 - 15.0 * t4;
  dw[1][2] = 1.5 * t2 - 4.0 * t3 + 2.5 * t4;
  d2w[0][0] = -60.0 * t + 180.0 * t2 - 120.0 * t3;
  d2w[0][1] = -36.0 * t + 96.0 * t2 - 60.0 * t3;
  d2w[0][2] = 1.0 - 9.0 * t + 18.0 * t2 - 10.0 * t3;
  d2w[1][0] = 60.0 * t - 180.0 * t2 + 120.0 * t3;
  d2w[1][1] = -24.0 * t + 84.0 * t2 - 60.0 * t3;
  d2w[1][2] = 3.0 * t - 12.0 * t2 + 10.0 * t3;
}

// s = fract(H^-1 pos) (src/mathkit/double3.cpp:22-45), the cell's lower corner and the position within the cell
__device__ __forceinline__ void locate(const InterpArgs& A, int q, size_t c0[3], double t[3])
{
  const double px = A.pos[3 * q], py = A.pos[3 * q + 1], pz = A.pos[3 * q + 2];
  for (int a = 0; a < 3; ++a)
  {
    double v = A.inv[a] * px + A.inv[3 + a] * py + A.inv[6 + a] * pz;
    v -= rint(v);
    if (v < 0.0) v += 1.0;
    if (v > 1.0) v -= 1.0;
    const size_t nc = size_t(A.n[a] - 1);
    size_t c = static_cast<size_t>(v * A.ncd[a]);
    c = c < nc - 1 ? c : nc - 1;
    c0[a] = c;
    t[a] = v * A.ncd[a] - double(c);
  }
}

// t-derivatives -> fractional (times the cell counts) -> Cartesian: dE/dr = H^-T dE/ds, d2E/dr2 = H^-T Hs H^-1
__device__ __forceinline__ void storeDerivatives(const InterpArgs& A, int q, double g0, double g1, double g2,
                                                 const double h[6] /* xx yy zz xy xz yz */)
{
  const double gs[3] = {g0 * A.ncd[0], g1 * A.ncd[1], g2 * A.ncd[2]};
  if (A.grad)
    for (int r = 0; r < 3; ++r) A.grad[3 * q + r] = A.inv[3 * r] * gs[0] + A.inv[3 * r + 1] * gs[1] + A.inv[3 * r + 2] * gs[2];
  if (A.hess)
  {
    double Hs[3][3];
    Hs[0][0] = h[0] * A.ncd[0] * A.ncd[0];
    Hs[1][1] = h[1] * A.ncd[1] * A.ncd[1];
    Hs[2][2] = h[2] * A.ncd[2] * A.ncd[2];
    Hs[0][1] = Hs[1][0] = h[3] * A.ncd[0] * A.ncd[1];
    Hs[0][2] = Hs[2][0] = h[4] * A.ncd[0] * A.ncd[2];
    Hs[1][2] = Hs[2][1] = h[5] * A.ncd[1] * A.ncd[2];
    // (H^-1)_{a i} = inv[3 i + a] (column-major cell inverse)
    double T[3][3];  // T[a][j] = sum_b Hs[a][b] (H^-1)_{b j}
    for (int a = 0; a < 3; ++a)
      for (int j = 0; j < 3; ++j) T[a][j] = Hs[a][0] * A.inv[3 * j] + Hs[a][1] * A.inv[3 * j + 1] + Hs[a][2] * A.inv[3 * j + 2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        A.hess[9 * q + 3 * i + j] = A.inv[3 * i] * T[0][j] + A.inv[3 * i + 1] * T[1][j] + A.inv[3 * i + 2] * T[2][j];
  }
}

// One corner's contribution: record F = {f, fx, fy, fz, fxy, fxz, fyz, fxyz}
__device__ __forceinline__ double corner(const double F[8], const double* x, const double* y, const double* z)
{
  const double a00 = fma(F[0], z[0], F[3] * z[1]);  // (dx,dy)=(0,0)
  const double a01 = fma(F[2], z[0], F[6] * z[1]);  // (0,1)
  const double a10 = fma(F[1], z[0], F[5] * z[1]);  // (1,0)
  const double a11 = fma(F[4], z[0], F[7] * z[1]);  // (1,1)
  return fma(x[0], fma(y[0], a00, y[1] * a01), x[1] * fma(y[0], a10, y[1] * a11));
}

__global__ __launch_bounds__(256) void k_interpolate(InterpArgs A)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= A.count) return;
  size_t c0[3];
  double t[3];
  locate(A, q, c0, t);
  double wx[2][2], wy[2][2], wz[2][2], dx[2][2], dy[2][2], dz[2][2], ex[2][2], ey[2][2], ez[2][2];
  hermite(t[0], wx, dx, ex);
  hermite(t[1], wy, dy, ey);
  hermite(t[2], wz, dz, ez);
  double e = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0, h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const bool wantGrad = A.grad != nullptr || A.hess != nullptr;
  const bool wantHess = A.hess != nullptr;
#pragma unroll
  for (int c = 0; c < 8; ++c)
  {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    const size_t idx = 8 * ((c0[0] + cx) + size_t(A.n[0]) * ((c0[1] + cy) + size_t(A.n[1]) * (c0[2] + cz)));
    const double2* p2 = reinterpret_cast<const double2*>(A.grid + idx);
    const double2 r0 = p2[0], r1 = p2[1], r2 = p2[2], r3 = p2[3];
    const double F[8] = {r0.x, r0.y, r1.x, r1.y, r2.x, r2.y, r3.x, r3.y};
    e += corner(F, wx[cx], wy[cy], wz[cz]);
    if (wantGrad)
    {
      g0 += corner(F, dx[cx], wy[cy], wz[cz]);
      g1 += corner(F, wx[cx], dy[cy], wz[cz]);
      g2 += corner(F, wx[cx], wy[cy], dz[cz]);
    }
    if (wantHess)
    {
      h[0] += corner(F, ex[cx], wy[cy], wz[cz]);
      h[1] += corner(F, wx[cx], ey[cy], wz[cz]);
      h[2] += corner(F, wx[cx], wy[cy], ez[cz]);
      h[3] += corner(F, dx[cx], dy[cy], wz[cz]);
      h[4] += corner(F, dx[cx], wy[cy], dz[cz]);
      h[5] += corner(F, wx[cx], dy[cy], dz[cz]);
    }
  }
  A.energy[q] = e;
  if (wantGrad) storeDerivatives(A, q, g0, g1, g2, h);
}

// Triquintic: value l of a record holds the derivative of orders kOrders[l]; its weight at corner (cx, cy, cz) is
// w_x[cx][o_x] w_y[cy][o_y] w_z[cz][o_z].
__global__ __launch_bounds__(256) void k_interpolate_q(InterpArgs A)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= A.count) return;
  size_t c0[3];
  double t[3];
  locate(A, q, c0, t);
  double w[3][2][3], dw[3][2][3], d2w[3][2][3];
  for (int a = 0; a < 3; ++a) hermite5(t[a], w[a], dw[a], d2w[a]);
  double e = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0, h[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const bool wantGrad = A.grad != nullptr || A.hess != nullptr;
  const bool wantHess = A.hess != nullptr;
  for (int c = 0; c < 8; ++c)
  {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
    const double* F = A.grid + 27 * ((c0[0] + cx) + size_t(A.n[0]) * ((c0[1] + cy) + size_t(A.n[1]) * (c0[2] + cz)));
    // this corner's 1-D weights, selected without indexing by the corner (which would move the bases to scratch)
    double X[3], Y[3], Z[3], DX[3], DY[3], DZ[3], EX[3], EY[3], EZ[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      X[d] = cx ? w[0][1][d] : w[0][0][d];
      Y[d] = cy ? w[1][1][d] : w[1][0][d];
      Z[d] = cz ? w[2][1][d] : w[2][0][d];
      DX[d] = cx ? dw[0][1][d] : dw[0][0][d];
      DY[d] = cy ? dw[1][1][d] : dw[1][0][d];
      DZ[d] = cz ? dw[2][1][d] : dw[2][0][d];
      EX[d] = cx ? d2w[0][1][d] : d2w[0][0][d];
      EY[d] = cy ? d2w[1][1][d] : d2w[1][0][d];
      EZ[d] = cz ? d2w[2][1][d] : d2w[2][0][d];
    }
#pragma unroll
    for (int l = 0; l < 27; ++l)
    {
      const int ox = rgq::kOrders[l][0], oy = rgq::kOrders[l][1], oz = rgq::kOrders[l][2];
      const double f = F[l];
      e = fma(f, X[ox] * Y[oy] * Z[oz], e);
      if (wantGrad)
      {
        g0 = fma(f, DX[ox] * Y[oy] * Z[oz], g0);
        g1 = fma(f, X[ox] * DY[oy] * Z[oz], g1);
        g2 = fma(f, X[ox] * Y[oy] * DZ[oz], g2);
        if (wantHess)
        {
          h[0] = fma(f, EX[ox] * Y[oy] * Z[oz], h[0]);
          h[1] = fma(f, X[ox] * EY[oy] * Z[oz], h[1]);
          h[2] = fma(f, X[ox] * Y[oy] * EZ[oz], h[2]);
          h[3] = fma(f, DX[ox] * DY[oy] * Z[oz], h[3]);
          h[4] = fma(f, DX[ox] * Y[oy] * DZ[oz], h[4]);
          h[5] = fma(f, X[ox] * DY[oy] * DZ[oz], h[5]);
        }
      }
    }
  }
  A.energy[q] = e;
  if (wantGrad) storeDerivatives(A, q, g0, g1, g2, h);
}

rg_status interpolateDevice(rg_ctx* ctx, const rg_grid* grid, int n, const double* d_pos, double* d_energy,
                            double* d_grad, double* d_hess)
{
  if (n <= 0) return RG_OK;
  InterpArgs A{};
  std::memcpy(A.inv, grid->inv, sizeof(A.inv));
  for (int a = 0; a < 3; ++a)
  {
    A.n[a] = grid->n[a];
    A.ncd[a] = double(grid->n[a] - 1);
  }
  A.grid = grid->d;
  A.pos = d_pos;
  A.count = n;
  A.energy = d_energy;
  A.grad = d_grad;
  A.hess = d_hess;
  hipLaunchKernelGGL(grid->values == 27 ? k_interpolate_q : k_interpolate, dim3((n + 255) / 256), dim3(256), 0,
                     ctx->stream, A);
  RG_HIP(ctx, hipGetLastError());
  return RG_OK;
}

// Host positions in, host energies (and derivatives) out, through one temporary device buffer.
rg_status interpolateHost(rg_ctx* ctx, const rg_grid* grid, int n, const double* pos, double* energy, double* gradient,
                          double* hessian, const char* who)
{
  if (!ctx || !grid || n < 0 || (n > 0 && (!pos || !energy)))
    return rg::fail(ctx, RG_ERR_INVALID, "%s: invalid arguments", who);
  if (n == 0) return RG_OK;
  RG_ENTER(ctx);
  double* d = nullptr;
  const size_t bytes = size_t(n) * (3 + 1 + (gradient ? 3 : 0) + (hessian ? 9 : 0)) * sizeof(double);
  RG_HIP(ctx, hipMalloc(&d, bytes));
  double* dpos = d;
  double* de = d + 3 * size_t(n);
  double* dg = gradient ? de + n : nullptr;
  double* dh = hessian ? de + n + (gradient ? 3 * size_t(n) : 0) : nullptr;
  rg_status st = RG_OK;
  hipError_t e = hipMemcpyAsync(dpos, pos, 3 * size_t(n) * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) st = rg::hipfail(ctx, e, "hipMemcpyAsync");
  if (st == RG_OK) st = rg::interpolateDevice(ctx, grid, n, dpos, de, dg, dh);
  if (st == RG_OK)
  {
    e = hipMemcpyAsync(energy, de, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && gradient)
      e = hipMemcpyAsync(gradient, dg, 3 * size_t(n) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && hessian)
      e = hipMemcpyAsync(hessian, dh, 9 * size_t(n) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) st = rg::hipfail(ctx, e, "interpolate copy-back");
  }
  hipStreamSynchronize(ctx->stream);
  hipFree(d);
  return st;
}

}  // namespace rg

extern "C" rg_status rg_grid_interpolate(rg_ctx* ctx, const rg_grid* grid, int n, const double* pos, double* energy,
                                         double* gradient)
{
  return rg::interpolateHost(ctx, grid, n, pos, energy, gradient, nullptr, "rg_grid_interpolate");
}

extern "C" rg_status rg_grid_interpolate_hessian(rg_ctx* ctx, const rg_grid* grid, int n, const double* pos,
                                                 double* energy, double* gradient, double* hessian)
{
  if (n > 0 && (!gradient || !hessian))
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_interpolate_hessian: gradient and hessian are required");
  return rg::interpolateHost(ctx, grid, n, pos, energy, gradient, hessian, "rg_grid_interpolate_hessian");
}

// Batched tricubic interpolation on gfx950.
//
// Replaces InterpolationEnergyGrid::interpolate / interpolateGradient, tricubic branch
// (src/raspakit/interpolation_energy_grid.cpp:3980-4070 and :4189-4347).  The reference gathers the 8 corner
// records X (64 values), forms the 64 polynomial coefficients a = A X with the fixed tricubic matrix
// (src/raspakit/interpolation_energy_grid.cpp:193-358) and sums a_ijk x^i y^j z^k.  That polynomial is the
// unique tricubic matching value and the seven mixed first derivatives at the 8 corners, which is the tensor
// product of 1-D cubic Hermite interpolants; this kernel evaluates it in that form (about 110 FMAs instead of a
// 64x64 product), so the result equals the reference's up to rounding.
//
// One lane per position; the 8 corner records are 64-byte contiguous chunks (component-fastest layout) read
// as four 16-byte loads each.
#include <cmath>
#include <cstring>

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
};

// Hermite basis on [0,1]: w[c][d] = weight of corner c for derivative order d
__device__ __forceinline__ void hermite(double t, double w[2][2], double dw[2][2])
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
  const double px = A.pos[3 * q], py = A.pos[3 * q + 1], pz = A.pos[3 * q + 2];
  // s = fract(H^-1 pos) (src/mathkit/double3.cpp:22-45)
  double s[3];
  for (int a = 0; a < 3; ++a)
  {
    double v = A.inv[a] * px + A.inv[3 + a] * py + A.inv[6 + a] * pz;
    v -= rint(v);
    if (v < 0.0) v += 1.0;
    if (v > 1.0) v -= 1.0;
    s[a] = v;
  }
  size_t c0[3];
  double t[3];
  for (int a = 0; a < 3; ++a)
  {
    const size_t nc = size_t(A.n[a] - 1);
    size_t c = static_cast<size_t>(s[a] * A.ncd[a]);
    c = c < nc - 1 ? c : nc - 1;
    c0[a] = c;
    t[a] = s[a] * A.ncd[a] - double(c);
  }
  double wx[2][2], wy[2][2], wz[2][2], dx[2][2], dy[2][2], dz[2][2];
  hermite(t[0], wx, dx);
  hermite(t[1], wy, dy);
  hermite(t[2], wz, dz);
  double e = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
  const bool wantGrad = A.grad != nullptr;
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
  }
  A.energy[q] = e;
  if (wantGrad)
  {
    // dE/ds_a = nc_a dE/dt_a ; dE/dr = H^-T dE/ds
    const double gs0 = g0 * A.ncd[0], gs1 = g1 * A.ncd[1], gs2 = g2 * A.ncd[2];
    for (int r = 0; r < 3; ++r) A.grad[3 * q + r] = A.inv[3 * r] * gs0 + A.inv[3 * r + 1] * gs1 + A.inv[3 * r + 2] * gs2;
  }
}

rg_status interpolateDevice(rg_ctx* ctx, const rg_grid* grid, int n, const double* d_pos, double* d_energy,
                            double* d_grad)
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
  hipLaunchKernelGGL(k_interpolate, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, A);
  RG_HIP(ctx, hipGetLastError());
  return RG_OK;
}

}  // namespace rg

extern "C" rg_status rg_grid_interpolate(rg_ctx* ctx, const rg_grid* grid, int n, const double* pos, double* energy,
                                         double* gradient)
{
  if (!ctx || !grid || n < 0 || (n > 0 && (!pos || !energy)))
    return rg::fail(ctx, RG_ERR_INVALID, "rg_grid_interpolate: invalid arguments");
  if (n == 0) return RG_OK;
  RG_ENTER(ctx);
  double* d = nullptr;
  const size_t bytes = size_t(n) * (3 + 1 + (gradient ? 3 : 0)) * sizeof(double);
  RG_HIP(ctx, hipMalloc(&d, bytes));
  double* dpos = d;
  double* de = d + 3 * size_t(n);
  double* dg = gradient ? de + n : nullptr;
  rg_status st = RG_OK;
  hipError_t e = hipMemcpyAsync(dpos, pos, 3 * size_t(n) * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) st = rg::hipfail(ctx, e, "hipMemcpyAsync");
  if (st == RG_OK) st = rg::interpolateDevice(ctx, grid, n, dpos, de, dg);
  if (st == RG_OK)
  {
    e = hipMemcpyAsync(energy, de, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && gradient)
      e = hipMemcpyAsync(gradient, dg, 3 * size_t(n) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) st = rg::hipfail(ctx, e, "interpolate copy-back");
  }
  hipStreamSynchronize(ctx->stream);
  hipFree(d);
  return st;
}

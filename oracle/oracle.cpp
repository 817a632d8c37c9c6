// FP64 CPU oracle for RASPA3's framework–adsorbate energy path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, in plain C++ and in double precision, the reference
// algorithms that the HIP library reimplements.  It is the checker: only the test-suite, the smoke
// check and the CPU-baseline leg of the benchmark may load it.  The product library never links it,
// never calls it and never falls back to it.
//
// Every function cites the reference file:line it follows (paths relative to the RASPA3 source
// tree).  Loop orders follow the reference so that the timed CPU baseline does the reference's work.
//
// Parity status per function (how each is pinned; see tests/test_oracle_golden.py):
//   units, LB mixing, shift, PBC, potentialVDW, potentialCoulomb  -> reference golden energies
//   Ewald Fourier total / single ion                              -> reference golden energies
//   Ewald Fourier difference                                      -> E(after) - E(before) of the pinned total
//   tricubic grid builder                                         -> brute-force image sums, closed form
//                                                                    single-atom values, finite differences
//   tricubic interpolation                                        -> exact reproduction of tricubic
//                                                                    polynomials, node values
//   trial selection                                               -> exact given the uniform
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace
{

// ---------------------------------------------------------------------------------------------------
// Units (src/foundationkit/units.ixx:25-60 constants; src/raspakit/units.cpp:15-100 derived values)
// ---------------------------------------------------------------------------------------------------
struct UnitsT
{
  double energyToKelvin, kelvinToEnergy, coulomb, kB;
  UnitsT()
  {
    const double L = 1e-10, T = 1e-12, M = 1.6605402e-27, Q = 1.60217733e-19;
    const double R = 8.314464919, NA = 6.0221419947e23, kB_SI = 1.380650324e-23, eps0 = 8.8541878176e-12;
    const double E = M * L * L / (T * T);
    energyToKelvin = E * NA / R;
    kelvinToEnergy = R / (E * NA);
    coulomb = Q * Q / (4.0 * M_PI * eps0 * L * E);
    kB = kB_SI / E;
  }
};
const UnitsT U;

// ---------------------------------------------------------------------------------------------------
// 3x3 matrices stored by columns: c[col][row]; columns are the lattice vectors a, b, c.
// src/mathkit/double3x3.cpp:102-118 (inverse), src/raspakit/simulationbox.cpp:40-60 (cell from lengths)
// ---------------------------------------------------------------------------------------------------
struct M3
{
  double c[3][3];
};

inline void mulv(const M3& m, const double v[3], double out[3])
{
  for (int r = 0; r < 3; ++r) out[r] = m.c[0][r] * v[0] + m.c[1][r] * v[1] + m.c[2][r] * v[2];
}

M3 inverse(const M3& a)
{
  // mRC = row R, column C
  const double m11 = a.c[0][0], m21 = a.c[0][1], m31 = a.c[0][2];
  const double m12 = a.c[1][0], m22 = a.c[1][1], m32 = a.c[1][2];
  const double m13 = a.c[2][0], m23 = a.c[2][1], m33 = a.c[2][2];
  double det = m11 * (m22 * m33 - m23 * m32) - m12 * (m21 * m33 - m23 * m31) + m13 * (m21 * m32 - m22 * m31);
  M3 inv;
  inv.c[0][0] = (m22 * m33 - m32 * m23) / det;
  inv.c[0][1] = -(m21 * m33 - m31 * m23) / det;
  inv.c[0][2] = (m21 * m32 - m31 * m22) / det;
  inv.c[1][0] = -(m12 * m33 - m32 * m13) / det;
  inv.c[1][1] = (m11 * m33 - m31 * m13) / det;
  inv.c[1][2] = -(m11 * m32 - m31 * m12) / det;
  inv.c[2][0] = (m12 * m23 - m22 * m13) / det;
  inv.c[2][1] = -(m11 * m23 - m21 * m13) / det;
  inv.c[2][2] = (m11 * m22 - m21 * m12) / det;
  return inv;
}

double determinant(const M3& a)
{
  const double m11 = a.c[0][0], m21 = a.c[0][1], m31 = a.c[0][2];
  const double m12 = a.c[1][0], m22 = a.c[1][1], m32 = a.c[1][2];
  const double m13 = a.c[2][0], m23 = a.c[2][1], m33 = a.c[2][2];
  return m11 * (m22 * m33 - m23 * m32) - m12 * (m21 * m33 - m23 * m31) + m13 * (m21 * m32 - m22 * m31);
}

struct Box
{
  M3 cell, inv;
  bool rect;
  double volume;
};

Box makeBox(const double* cell9, int rect)
{
  Box b;
  for (int col = 0; col < 3; ++col)
    for (int r = 0; r < 3; ++r) b.cell.c[col][r] = cell9[3 * col + r];
  b.inv = inverse(b.cell);
  b.volume = determinant(b.cell);
  b.rect = rect != 0;
  return b;
}

// src/raspakit/simulationbox.ixx:131-161
inline void applyPBC(const Box& b, double dr[3])
{
  if (b.rect)
  {
    for (int a = 0; a < 3; ++a)
    {
      double t = dr[a] * b.inv.c[a][a] + ((dr[a] >= 0.0) ? 0.5 : -0.5);
      dr[a] = dr[a] - static_cast<double>(static_cast<std::int64_t>(t)) * b.cell.c[a][a];
    }
  }
  else
  {
    double s[3];
    mulv(b.inv, dr, s);
    for (int a = 0; a < 3; ++a) s[a] -= static_cast<double>(static_cast<std::int64_t>(s[a] + ((s[a] >= 0.0) ? 0.5 : -0.5)));
    mulv(b.cell, s, dr);
  }
}

// src/raspakit/simulationbox.cpp:137-148
void perpendicularWidths(const M3& m, double w[3])
{
  auto cross = [](const double* x, const double* y, double* o)
  {
    o[0] = x[1] * y[2] - x[2] * y[1];
    o[1] = x[2] * y[0] - x[0] * y[2];
    o[2] = x[0] * y[1] - x[1] * y[0];
  };
  double c1[3], c2[3], c3[3];
  cross(m.c[0], m.c[1], c1);
  cross(m.c[1], m.c[2], c2);
  cross(m.c[0], m.c[2], c3);
  double v = std::fabs(m.c[0][0] * c2[0] + m.c[0][1] * c2[1] + m.c[0][2] * c2[2]);
  auto dot = [](const double* x) { return x[0] * x[0] + x[1] * x[1] + x[2] * x[2]; };
  w[0] = v / std::sqrt(dot(c2));
  w[1] = v / std::sqrt(dot(c3));
  w[2] = v / std::sqrt(dot(c1));
}

// ---------------------------------------------------------------------------------------------------
// Pair table (src/raspakit/forcefield/forcefield.cpp:104-131 ctor, :924-1014 mixing, :1038-1051 shift;
// src/raspakit/forcefield/vdwparameters.ixx:113-118 K->internal, :206-210 shift,
// src/raspakit/forcefield/vdwparameters.cpp:211-223 LJ energy at full coupling)
// ---------------------------------------------------------------------------------------------------
struct Pair
{
  double eps;    // internal energy units
  double sigma;  // Angstrom
  double shift;  // internal energy units
  int lj;        // 1 = Lennard-Jones, 0 = None
};

struct FF
{
  int n;
  const double* eps;
  const double* sigma;
  const double* shift;
  const int* lj;
  Pair operator()(int i, int j) const
  {
    int k = i * n + j;
    return {eps[k], sigma[k], shift[k], lj[k]};
  }
};

// ---------------------------------------------------------------------------------------------------
// Tricubic radial factors (src/raspakit/potentials/interpolation/tricubic_derivative_lj.ixx:33-46,
// tricubic_derivative_real_ewald.ixx:32-56)
// ---------------------------------------------------------------------------------------------------
struct Factor
{
  double e, f1, f2, f3;
};

inline Factor ljTricubic(const Pair& p, double r2)
{
  double arg1 = 4.0 * p.eps;
  double arg2 = p.sigma * p.sigma;
  double arg3 = p.shift;
  double temp3 = (arg2 / r2) * (arg2 / r2) * (arg2 / r2);
  double r4 = r2 * r2;
  double r6 = r4 * r2;
  return {arg1 * (temp3 * temp3 - temp3) - arg3, -6.0 * arg1 * (2.0 * temp3 * temp3 - temp3) / r2,
          24.0 * arg1 * (7.0 * temp3 * temp3 - 2.0 * temp3) / r4,
          -96.0 * arg1 * (28.0 * temp3 * temp3 - 5.0 * temp3) / r6};
}

inline Factor ewaldTricubic(double alpha, double rr, double r, double chargeA, double chargeB)
{
  const double C = U.coulomb;
  double exp_factor = std::erfc(alpha * r);
  double temp = C * chargeA * chargeB * exp_factor / r;
  double exp_term_rr = 2.0 * alpha * r * std::exp(-alpha * alpha * rr) * (1.0 / std::sqrt(M_PI));
  double alpha_r = alpha * r;
  double alpha_r_2 = alpha_r * alpha_r;
  double alpha_r_4 = alpha_r_2 * alpha_r_2;
  double r3 = rr * r;
  double r5 = r3 * rr;
  double r7 = r5 * rr;
  return {temp, -C * chargeA * chargeB * ((exp_term_rr + exp_factor) / r3),
          C * chargeA * chargeB * ((exp_term_rr * (2.0 * alpha_r_2 + 3.0) + 3.0 * exp_factor) / r5),
          -C * chargeA * chargeB * ((exp_term_rr * (4.0 * alpha_r_4 + 10 * alpha_r_2 + 15.0) + 15.0 * exp_factor) / r7)};
}

struct AtomsView
{
  int n;
  const double* pos;  // n*3
  const int* type;
  const double* charge;
};

// src/raspakit/framework.cpp:1085-1103 (makeSuperCell)
void makeSuperCell(const M3& cell, const AtomsView& a, const int nc[3], std::vector<double>& pos, std::vector<int>& type,
                   std::vector<double>& charge)
{
  for (int i = 0; i < nc[0]; ++i)
    for (int j = 0; j < nc[1]; ++j)
      for (int k = 0; k < nc[2]; ++k)
      {
        double t[3] = {double(i), double(j), double(k)}, d[3];
        mulv(cell, t, d);
        for (int m = 0; m < a.n; ++m)
        {
          pos.push_back(a.pos[3 * m + 0] + d[0]);
          pos.push_back(a.pos[3 * m + 1] + d[1]);
          pos.push_back(a.pos[3 * m + 2] + d[2]);
          type.push_back(a.type[m]);
          charge.push_back(a.charge[m]);
        }
      }
}

// src/raspakit/interactions/framework_molecule_grid.cpp:168-282 (+ the triclinic helper :76-144)
void tricubicFractionalAtPosition(int gridType, const FF& ff, double rcVDW, double alpha, const Box& superBox,
                                  const double posA[3], int typeA, const Box& unitBox, int n, const double* pos,
                                  const int* type, const double* charge, double out[8])
{
  const double rc2 = rcVDW * rcVDW;
  const M3& fc = unitBox.cell;
  auto factor = [&](double rr, int typeB, double chargeB) -> Factor
  {
    if (gridType == 0) return ljTricubic(ff(typeA, typeB), rr);
    return ewaldTricubic(alpha, rr, std::sqrt(rr), 1.0, chargeB);
  };
  if (unitBox.rect)
  {
    double e = 0, fx = 0, fy = 0, fz = 0, sxy = 0, sxz = 0, syz = 0, txyz = 0;
    for (int m = 0; m < n; ++m)
    {
      double dr[3] = {posA[0] - pos[3 * m], posA[1] - pos[3 * m + 1], posA[2] - pos[3 * m + 2]};
      applyPBC(superBox, dr);
      double rr = std::max(0.1, dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      if (rr < rc2)
      {
        Factor v = factor(rr, type[m], charge[m]);
        e += v.e;
        fx += dr[0] * v.f1;
        fy += dr[1] * v.f1;
        fz += dr[2] * v.f1;
        sxy += v.f2 * dr[0] * dr[1];
        sxz += v.f2 * dr[0] * dr[2];
        syz += v.f2 * dr[1] * dr[2];
        txyz += v.f3 * dr[0] * dr[1] * dr[2];
      }
    }
    out[0] = e;
    out[1] = fx * fc.c[0][0];
    out[2] = fy * fc.c[1][1];
    out[3] = fz * fc.c[2][2];
    out[4] = sxy * fc.c[0][0] * fc.c[1][1];
    out[5] = sxz * fc.c[0][0] * fc.c[2][2];
    out[6] = syz * fc.c[1][1] * fc.c[2][2];
    out[7] = txyz * fc.c[0][0] * fc.c[1][1] * fc.c[2][2];
    return;
  }
  double e = 0, d1[3] = {0, 0, 0}, d2[3][3] = {}, d3[3][3][3] = {};
  for (int m = 0; m < n; ++m)
  {
    double dr[3] = {posA[0] - pos[3 * m], posA[1] - pos[3 * m + 1], posA[2] - pos[3 * m + 2]};
    applyPBC(superBox, dr);
    double rr = std::max(0.1, dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    if (rr < rc2)
    {
      Factor v = factor(rr, type[m], charge[m]);
      e += v.e;
      for (int i = 0; i < 3; ++i)
      {
        d1[i] += dr[i] * v.f1;
        for (int j = 0; j < 3; ++j)
        {
          d2[i][j] += v.f2 * dr[i] * dr[j] + (i == j ? v.f1 : 0.0);
          for (int k = 0; k < 3; ++k)
            d3[i][j][k] += v.f3 * dr[i] * dr[j] * dr[k] + (j == k ? v.f2 * dr[i] : 0.0) + (i == k ? v.f2 * dr[j] : 0.0) +
                           (i == j ? v.f2 * dr[k] : 0.0);
        }
      }
    }
  }
  double f1[3] = {0, 0, 0}, f01 = 0, f02 = 0, f12 = 0, f012 = 0;
  for (int p = 0; p < 3; ++p)
  {
    f1[0] += fc.c[0][p] * d1[p];
    f1[1] += fc.c[1][p] * d1[p];
    f1[2] += fc.c[2][p] * d1[p];
    for (int q = 0; q < 3; ++q)
    {
      f01 += fc.c[0][p] * fc.c[1][q] * d2[p][q];
      f02 += fc.c[0][p] * fc.c[2][q] * d2[p][q];
      f12 += fc.c[1][p] * fc.c[2][q] * d2[p][q];
      for (int r = 0; r < 3; ++r) f012 += fc.c[0][p] * fc.c[1][q] * fc.c[2][r] * d3[p][q][r];
    }
  }
  out[0] = e;
  out[1] = f1[0];
  out[2] = f1[1];
  out[3] = f1[2];
  out[4] = f01;
  out[5] = f02;
  out[6] = f12;
  out[7] = f012;
}

// ---------------------------------------------------------------------------------------------------
// Tricubic Hermite coefficient map.  The reference applies a fixed 64x64 integer matrix
// (src/raspakit/interpolation_energy_grid.cpp:55 table, :193-358 unrolled product).  The oracle builds
// that matrix from its definition: B[8d+c][i+4j+16k] is the (dx,dy,dz)-derivative of x^i y^j z^k at
// corner c, and the coefficient map is B^-1 (Lekien & Marsden, IJNME 63:455, 2005).
// ---------------------------------------------------------------------------------------------------
struct TricubicMatrix
{
  double A[64][64];
  TricubicMatrix()
  {
    static const int D[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    double B[64][128];
    auto dmono = [](int p, int d, double x) -> double
    {
      // d-th derivative (d in {0,1}) of x^p at x in {0,1}
      if (d == 0) return (p == 0) ? 1.0 : std::pow(x, p);
      if (p == 0) return 0.0;
      return p * ((p - 1 == 0) ? 1.0 : std::pow(x, p - 1));
    };
    for (int d = 0; d < 8; ++d)
      for (int c = 0; c < 8; ++c)
      {
        double x = c & 1, y = (c >> 1) & 1, z = (c >> 2) & 1;
        for (int k = 0; k < 4; ++k)
          for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i)
              B[8 * d + c][i + 4 * j + 16 * k] = dmono(i, D[d][0], x) * dmono(j, D[d][1], y) * dmono(k, D[d][2], z);
        for (int m = 0; m < 64; ++m) B[8 * d + c][64 + m] = (m == 8 * d + c) ? 1.0 : 0.0;
      }
    // Gauss-Jordan with partial pivoting
    for (int col = 0; col < 64; ++col)
    {
      int piv = col;
      for (int r = col + 1; r < 64; ++r)
        if (std::fabs(B[r][col]) > std::fabs(B[piv][col])) piv = r;
      for (int m = 0; m < 128; ++m) std::swap(B[col][m], B[piv][m]);
      double inv = 1.0 / B[col][col];
      for (int m = 0; m < 128; ++m) B[col][m] *= inv;
      for (int r = 0; r < 64; ++r)
        if (r != col && B[r][col] != 0.0)
        {
          double f = B[r][col];
          for (int m = 0; m < 128; ++m) B[r][m] -= f * B[col][m];
        }
    }
    // The exact inverse has integer entries; round away the elimination noise.
    for (int r = 0; r < 64; ++r)
      for (int m = 0; m < 64; ++m) A[r][m] = std::nearbyint(B[r][64 + m]);
  }
};
const TricubicMatrix TM;

// src/raspakit/interpolation_energy_grid.cpp:3980-4070 (tricubic branch of interpolate) and :4189-4347
// (gradient): s = fract(H^-1 pos), cell index = min(floor(s*nc), nc-1), gather, a = A X, polynomial.
void interpolateOne(const Box& unit, const int np[3], const double* data, const double pos[3], double* e,
                    double grad[3], double* hess = nullptr)
{
  const int nc[3] = {np[0] - 1, np[1] - 1, np[2] - 1};
  double s[3];
  mulv(unit.inv, pos, s);
  for (int a = 0; a < 3; ++a)
  {
    s[a] -= std::rint(s[a]);
    if (s[a] < 0.0) s[a] += 1.0;
    if (s[a] > 1.0) s[a] -= 1.0;
  }
  std::size_t x0 = std::min(static_cast<std::size_t>(s[0] * nc[0]), static_cast<std::size_t>(nc[0] - 1));
  std::size_t y0 = std::min(static_cast<std::size_t>(s[1] * nc[1]), static_cast<std::size_t>(nc[1] - 1));
  std::size_t z0 = std::min(static_cast<std::size_t>(s[2] * nc[2]), static_cast<std::size_t>(nc[2] - 1));
  double t[3] = {s[0] * nc[0] - double(x0), s[1] * nc[1] - double(y0), s[2] * nc[2] - double(z0)};
  double X[64], a[64];
  auto at = [&](int l, std::size_t i, std::size_t j, std::size_t k)
  { return data[l + 8 * (i + std::size_t(np[0]) * (j + std::size_t(np[1]) * k))]; };
  for (int l = 0; l < 8; ++l)
  {
    X[8 * l + 0] = at(l, x0, y0, z0);
    X[8 * l + 1] = at(l, x0 + 1, y0, z0);
    X[8 * l + 2] = at(l, x0, y0 + 1, z0);
    X[8 * l + 3] = at(l, x0 + 1, y0 + 1, z0);
    X[8 * l + 4] = at(l, x0, y0, z0 + 1);
    X[8 * l + 5] = at(l, x0 + 1, y0, z0 + 1);
    X[8 * l + 6] = at(l, x0, y0 + 1, z0 + 1);
    X[8 * l + 7] = at(l, x0 + 1, y0 + 1, z0 + 1);
  }
  for (int r = 0; r < 64; ++r)
  {
    double acc = 0.0;
    for (int m = 0; m < 64; ++m) acc += TM.A[r][m] * X[m];
    a[r] = acc;
  }
  double v = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
  double H[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  // p(x, n, d): d-th derivative of x^n
  auto p = [](double x, int n, int d)
  {
    if (d > n) return 0.0;
    double f = 1.0;
    for (int m = 0; m < d; ++m) f *= double(n - m);
    return f * std::pow(x, n - d);
  };
  for (int k = 0; k < 4; ++k)
    for (int j = 0; j < 4; ++j)
      for (int i = 0; i < 4; ++i)
      {
        double c = a[i + 4 * j + 16 * k];
        double px = std::pow(t[0], i), py = std::pow(t[1], j), pz = std::pow(t[2], k);
        v += c * px * py * pz;
        if (grad)
        {
          if (i > 0) gx += i * c * std::pow(t[0], i - 1) * py * pz;
          if (j > 0) gy += j * c * px * std::pow(t[1], j - 1) * pz;
          if (k > 0) gz += k * c * px * py * std::pow(t[2], k - 1);
        }
        if (hess)
        {
          const int ijk[3] = {i, j, k};
          for (int a1 = 0; a1 < 3; ++a1)
            for (int a2 = 0; a2 < 3; ++a2)
            {
              double term = c;
              for (int ax = 0; ax < 3; ++ax) term *= p(t[ax], ijk[ax], int(ax == a1) + int(ax == a2));
              H[a1][a2] += term;
            }
        }
      }
  *e = v;
  if (hess)
  {
    // d2E/dr_i dr_j = sum_ab (H^-1)_ai nc_a nc_b d2E/dt_a dt_b (H^-1)_bj
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
      {
        double acc = 0.0;
        for (int a1 = 0; a1 < 3; ++a1)
          for (int a2 = 0; a2 < 3; ++a2)
            acc += unit.inv.c[i][a1] * H[a1][a2] * double(nc[a1]) * double(nc[a2]) * unit.inv.c[j][a2];
        hess[3 * i + j] = acc;
      }
  }
  if (grad)
  {
    // d/ds = nc * d/dt ; d/dr = H^-T d/ds
    double gs[3] = {gx * nc[0], gy * nc[1], gz * nc[2]};
    for (int r = 0; r < 3; ++r) grad[r] = unit.inv.c[r][0] * gs[0] + unit.inv.c[r][1] * gs[1] + unit.inv.c[r][2] * gs[2];
  }
}

// ---------------------------------------------------------------------------------------------------
// Pair energies for trial sites (src/raspakit/potentials/pair_vdw.ixx:74-111 LJ branch, Order 0;
// src/raspakit/potentials/pair_coulomb.ixx:39-76 Ewald branch, Order 0; offset delta = 0.01,
// src/raspakit/potentials/coulomb_real_space.ixx:25)
// ---------------------------------------------------------------------------------------------------
inline double vdwEnergy(const Pair& p, double scaling, double rr)
{
  if (!p.lj) return 0.0;
  double arg1 = 4.0 * p.eps;
  double arg2 = p.sigma * p.sigma;
  double arg3 = p.shift;
  double temp = rr / arg2;
  double temp3 = temp * temp * temp;
  double inv_scaling = 1.0 - scaling;
  double rri3 = 1.0 / (temp3 + 0.5 * inv_scaling * inv_scaling);
  double term = arg1 * (rri3 * (rri3 - 1.0)) - arg3;
  return scaling * term;
}

inline double coulombEnergy(double alpha, double scaling, double r, double qA, double qB)
{
  double prefactor = U.coulomb * qA * qB;
  double s = r + 0.01 * (1.0 - scaling);
  double phi = std::erfc(alpha * s) / s;
  return scaling * prefactor * phi;
}

struct Site
{
  double pos[3];
  int type;
  double charge, scalingVDW, scalingCoulomb;
  int molecule;
};

}  // namespace

#include "oracle_quintic.inc"

// ===================================================================================================
// C entry points (ctypes)
// ===================================================================================================
extern "C"
{

void orc_units(double* out)
{
  out[0] = U.energyToKelvin;
  out[1] = U.kelvinToEnergy;
  out[2] = U.coulomb;
  out[3] = U.kB;
}

// SimulationBox(a, b, c, alpha, beta, gamma) -- src/raspakit/simulationbox.cpp:40-60.  Angles in
// degrees are converted as pi * deg / 180 (the reference's own conversion).  If 'lengths_only' is set the
// (a, b, c) constructor is followed instead (exactly diagonal, src/raspakit/simulationbox.cpp:15-37).
void orc_box(double a, double b, double c, double alpha_deg, double beta_deg, double gamma_deg, int lengths_only,
             double* cell9, double* inv9, int* rect, double* volume, double* widths3)
{
  M3 m{};
  int r = 1;
  if (lengths_only)
  {
    m.c[0][0] = a;
    m.c[1][1] = b;
    m.c[2][2] = c;
  }
  else
  {
    double al = M_PI * alpha_deg / 180.0, be = M_PI * beta_deg / 180.0, ga = M_PI * gamma_deg / 180.0;
    r = (std::fabs(al - 0.5 * M_PI) < 1e-5) && (std::fabs(be - 0.5 * M_PI) < 1e-5) && (std::fabs(ga - 0.5 * M_PI) < 1e-5);
    double temp = (std::cos(al) - std::cos(ga) * std::cos(be)) / std::sin(ga);
    m.c[0][0] = a;
    m.c[1][0] = b * std::cos(ga);
    m.c[1][1] = b * std::sin(ga);
    m.c[2][0] = c * std::cos(be);
    m.c[2][1] = c * temp;
    m.c[2][2] = c * std::sqrt(1.0 - std::cos(be) * std::cos(be) - temp * temp);
  }
  M3 inv = inverse(m);
  for (int col = 0; col < 3; ++col)
    for (int rr = 0; rr < 3; ++rr)
    {
      cell9[3 * col + rr] = m.c[col][rr];
      inv9[3 * col + rr] = inv.c[col][rr];
    }
  *rect = r;
  *volume = determinant(m);
  perpendicularWidths(m, widths3);
}

// ForceField pair table: LB mixing, None-typing, shift at the cutoff (pair uses the framework cutoff if
// either pseudo-atom is a framework atom, src/raspakit/forcefield/forcefield.cpp:1016-1024).
void orc_ff_pairs(int n, const double* eps_K, const double* sigma, const int* is_framework, int shifted,
                  double rc_framework, double rc_molecule, double* out_eps, double* out_sigma, double* out_shift,
                  int* out_lj)
{
  for (int i = 0; i < n * n; ++i)
  {
    out_eps[i] = 0.0;
    out_sigma[i] = 0.0;
    out_shift[i] = 0.0;
    out_lj[i] = 1;
  }
  for (int i = 0; i < n; ++i)
  {
    out_eps[i * n + i] = eps_K[i] * U.kelvinToEnergy;
    out_sigma[i * n + i] = sigma[i];
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
    {
      double mix0 = std::sqrt(out_eps[i * n + i] * out_eps[j * n + j]);
      double mix1 = 0.5 * (out_sigma[i * n + i] + out_sigma[j * n + j]);
      out_eps[i * n + j] = out_eps[j * n + i] = mix0;
      out_sigma[i * n + j] = out_sigma[j * n + i] = mix1;
    }
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j)
    {
      bool zi = out_eps[i * n + i] == 0.0 || out_sigma[i * n + i] == 0.0;
      bool zj = out_eps[j * n + j] == 0.0 || out_sigma[j * n + j] == 0.0;
      if (zi || zj) out_lj[i * n + j] = out_lj[j * n + i] = 0;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
    {
      int k = i * n + j;
      if (shifted && out_lj[k])
      {
        double rc = (is_framework[i] || is_framework[j]) ? rc_framework : rc_molecule;
        double rr = rc * rc;
        double arg2 = out_sigma[k] * out_sigma[k];
        double temp = rr / arg2;
        double rri3 = 1.0 / (temp * temp * temp);
        out_shift[k] = 4.0 * out_eps[k] * (rri3 * (rri3 - 1.0));
      }
    }
}

// Automatic Ewald parameters (src/raspakit/forcefield/forcefield.cpp:1415-1438).
void orc_ewald_auto(double precision, double rc_coulomb, const double* widths3, double* alpha, int* kmax3)
{
  double eps = std::min(std::fabs(precision), 0.5);
  double tol = std::sqrt(std::abs(std::log(eps * rc_coulomb)));
  *alpha = std::sqrt(std::abs(std::log(eps * rc_coulomb * tol))) / rc_coulomb;
  double tol1 = std::sqrt(-std::log(eps * rc_coulomb * (2.0 * tol * *alpha) * (2.0 * tol * *alpha)));
  for (int a = 0; a < 3; ++a) kmax3[a] = static_cast<int>(std::rint(0.25 + widths3[a] * *alpha * tol1 / M_PI));
}

// Grid dimensions n = (size_t)(w / h + 0.5), src/raspakit/system_interpolation_grids.cpp:192-198, 222-226.
void orc_grid_dims(const double* cell9, double spacing, int* n3)
{
  M3 m{};
  for (int col = 0; col < 3; ++col)
    for (int r = 0; r < 3; ++r) m.c[col][r] = cell9[3 * col + r];
  double w[3];
  perpendicularWidths(m, w);
  for (int a = 0; a < 3; ++a) n3[a] = static_cast<int>(static_cast<std::size_t>(w[a] / spacing + 0.5));
}

// InterpolationEnergyGrid::makeFrameworkInterpolationGrid, tricubic branch
// (src/raspakit/interpolation_energy_grid.cpp:3800-3875), restricted to z-planes [k_begin, k_end) so that a
// deterministic slab can be timed.  grid_type: 0 = LennardJones (probe pseudo-atom type_a), 3 = EwaldReal
// (unit-charge probe).  The supercell is sized from 'rc_supercell' (the caller passes the VDW cutoff for VDW
// grids and the Coulomb cutoff for the Coulomb grid, system_interpolation_grids.cpp:211-212, 234-235), pair
// inclusion always uses rc_vdw (framework_molecule_grid.cpp:173).  nthreads > 1 splits the z-planes.
int orc_grid_build(const double* cell9, int rect, int nx, int ny, int nz, int grid_type, int natoms,
                   const double* pos, const int* type, const double* charge, int ntypes, const double* pair_eps,
                   const double* pair_sigma, const double* pair_shift, const int* pair_lj, int type_a, double rc_vdw,
                   double rc_supercell, double alpha, double overlap, int k_begin, int k_end, int nthreads,
                   double* out)
{
  if (nx < 2 || ny < 2 || nz < 2) return -1;
  Box unit = makeBox(cell9, rect);
  double w[3];
  perpendicularWidths(unit.cell, w);
  int nc[3];
  for (int a = 0; a < 3; ++a) nc[a] = static_cast<int>(std::ceil(2.0 * rc_supercell / w[a]));
  std::vector<double> spos;
  std::vector<int> stype;
  std::vector<double> sq;
  AtomsView av{natoms, pos, type, charge};
  makeSuperCell(unit.cell, av, nc, spos, stype, sq);
  // scaled(int3) keeps the type and rebuilds the cell from lengths and angles; for the cells used here the
  // product of the unit cell by the integer factors is the same matrix up to rounding of cos(pi/2).
  Box super = unit;
  for (int col = 0; col < 3; ++col)
    for (int r = 0; r < 3; ++r) super.cell.c[col][r] = unit.cell.c[col][r] * nc[col];
  super.inv = inverse(super.cell);
  super.volume = determinant(super.cell);
  FF ff{ntypes, pair_eps, pair_sigma, pair_shift, pair_lj};
  const double delta[3] = {1.0 / double(nx - 1), 1.0 / double(ny - 1), 1.0 / double(nz - 1)};
  const int ns = static_cast<int>(stype.size());
  // rows (j, k) of planes k_begin .. k_end - 1, in the reference's k -> j -> i order within each worker's range
  auto work = [&](long r0, long r1)
  {
    for (long row = r0; row < r1; ++row)
    {
      const int k = k_begin + static_cast<int>(row / ny), j = static_cast<int>(row % ny);
        for (int i = 0; i < nx; ++i)
        {
          double s[3] = {double(i) * delta[0], double(j) * delta[1], double(k) * delta[2]}, p[3];
          mulv(unit.cell, s, p);
          double v[8];
          tricubicFractionalAtPosition(grid_type, ff, rc_vdw, alpha, super, p, type_a, unit, ns, spos.data(),
                                       stype.data(), sq.data(), v);
          v[1] /= double(nx - 1);
          v[2] /= double(ny - 1);
          v[3] /= double(nz - 1);
          v[4] /= double(nx - 1) * double(ny - 1);
          v[5] /= double(nx - 1) * double(nz - 1);
          v[6] /= double(ny - 1) * double(nz - 1);
          v[7] /= double(nx - 1) * double(ny - 1) * double(nz - 1);
          // out holds planes k_begin .. k_end - 1 only
          double* o = out + 8 * (std::size_t(i) + std::size_t(nx) * (std::size_t(j) + std::size_t(ny) * (k - k_begin)));
          if (v[0] > overlap)
          {
            o[0] = 2.0 * overlap;
            for (int l = 1; l < 4; ++l) o[l] = std::clamp(v[l], -overlap, overlap);
            for (int l = 4; l < 8; ++l) o[l] = 0.0;
          }
          else
            for (int l = 0; l < 8; ++l) o[l] = v[l];
        }
    }
  };
  const long rows = long(k_end - k_begin) * ny;
  if (nthreads <= 1)
    work(0, rows);
  else
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t)
    {
      const long a = (rows * t) / nthreads, b = (rows * (t + 1)) / nthreads;
      if (a < b) th.emplace_back(work, a, b);
    }
    for (auto& x : th) x.join();
  }
  return 0;
}

// Single grid point (no supercell loop order), for spot checks of a GPU grid at arbitrary (i, j, k).
int orc_grid_points(const double* cell9, int rect, int nx, int ny, int nz, int grid_type, int natoms,
                    const double* pos, const int* type, const double* charge, int ntypes, const double* pair_eps,
                    const double* pair_sigma, const double* pair_shift, const int* pair_lj, int type_a,
                    double rc_vdw, double rc_supercell, double alpha, double overlap, int npoints, const int* ijk,
                    int nthreads, double* out)
{
  Box unit = makeBox(cell9, rect);
  double w[3];
  perpendicularWidths(unit.cell, w);
  int nc[3];
  for (int a = 0; a < 3; ++a) nc[a] = static_cast<int>(std::ceil(2.0 * rc_supercell / w[a]));
  std::vector<double> spos;
  std::vector<int> stype;
  std::vector<double> sq;
  AtomsView av{natoms, pos, type, charge};
  makeSuperCell(unit.cell, av, nc, spos, stype, sq);
  Box super = unit;
  for (int col = 0; col < 3; ++col)
    for (int r = 0; r < 3; ++r) super.cell.c[col][r] = unit.cell.c[col][r] * nc[col];
  super.inv = inverse(super.cell);
  FF ff{ntypes, pair_eps, pair_sigma, pair_shift, pair_lj};
  const double delta[3] = {1.0 / double(nx - 1), 1.0 / double(ny - 1), 1.0 / double(nz - 1)};
  const int ns = static_cast<int>(stype.size());
  auto work = [&](int p0, int p1)
  {
    for (int q = p0; q < p1; ++q)
    {
      int i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
      double s[3] = {double(i) * delta[0], double(j) * delta[1], double(k) * delta[2]}, p[3];
      mulv(unit.cell, s, p);
      double v[8];
      tricubicFractionalAtPosition(grid_type, ff, rc_vdw, alpha, super, p, type_a, unit, ns, spos.data(),
                                   stype.data(), sq.data(), v);
      v[1] /= double(nx - 1);
      v[2] /= double(ny - 1);
      v[3] /= double(nz - 1);
      v[4] /= double(nx - 1) * double(ny - 1);
      v[5] /= double(nx - 1) * double(nz - 1);
      v[6] /= double(ny - 1) * double(nz - 1);
      v[7] /= double(nx - 1) * double(ny - 1) * double(nz - 1);
      double* o = out + 8 * std::size_t(q);
      if (v[0] > overlap)
      {
        o[0] = 2.0 * overlap;
        for (int l = 1; l < 4; ++l) o[l] = std::clamp(v[l], -overlap, overlap);
        for (int l = 4; l < 8; ++l) o[l] = 0.0;
      }
      else
        for (int l = 0; l < 8; ++l) o[l] = v[l];
    }
  };
  if (nthreads <= 1)
    work(0, npoints);
  else
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t)
    {
      int a = (npoints * t) / nthreads, b = (npoints * (t + 1)) / nthreads;
      if (a < b) th.emplace_back(work, a, b);
    }
    for (auto& x : th) x.join();
  }
  return 0;
}

// Per-point analytic energy only (Interactions::calculateEnergyAtPosition,
// src/raspakit/interactions/framework_molecule_grid.cpp:30-72), used by the grid validation statistics.
double orc_energy_at_position(const double* cell9, int rect, int grid_type, int natoms, const double* pos,
                              const int* type, const double* charge, int ntypes, const double* pair_eps,
                              const double* pair_sigma, const double* pair_shift, const int* pair_lj, int type_a,
                              double rc_vdw, double rc_supercell, double alpha, const double* p)
{
  Box unit = makeBox(cell9, rect);
  double w[3];
  perpendicularWidths(unit.cell, w);
  int nc[3];
  for (int a = 0; a < 3; ++a) nc[a] = static_cast<int>(std::ceil(2.0 * rc_supercell / w[a]));
  std::vector<double> spos;
  std::vector<int> stype;
  std::vector<double> sq;
  AtomsView av{natoms, pos, type, charge};
  makeSuperCell(unit.cell, av, nc, spos, stype, sq);
  Box super = unit;
  for (int col = 0; col < 3; ++col)
    for (int r = 0; r < 3; ++r) super.cell.c[col][r] = unit.cell.c[col][r] * nc[col];
  super.inv = inverse(super.cell);
  FF ff{ntypes, pair_eps, pair_sigma, pair_shift, pair_lj};
  double e = 0.0;
  for (std::size_t m = 0; m < stype.size(); ++m)
  {
    double dr[3] = {p[0] - spos[3 * m], p[1] - spos[3 * m + 1], p[2] - spos[3 * m + 2]};
    applyPBC(super, dr);
    double rr = std::max(0.1, dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    if (rr < rc_vdw * rc_vdw)
    {
      if (grid_type == 0)
        e += ljTricubic(ff(type_a, stype[m]), rr).e;
      else
        e += ewaldTricubic(alpha, rr, std::sqrt(rr), 1.0, sq[m]).e;
    }
  }
  return e;
}

// InterpolationEnergyGrid::interpolate / interpolateGradient (tricubic).  grad may be null.
void orc_interpolate(const double* cell9, int rect, int nx, int ny, int nz, const double* grid, int n,
                     const double* pos, double* e, double* grad)
{
  Box unit = makeBox(cell9, rect);
  const int np[3] = {nx, ny, nz};
  for (int q = 0; q < n; ++q) interpolateOne(unit, np, grid, pos + 3 * q, e + q, grad ? grad + 3 * q : nullptr);
}

// Tricubic energies, gradients and Hessians (n*9, row-major, Cartesian) of the interpolant.  The reference's
// interpolateHessian returns a zero Hessian for this scheme (src/raspakit/interpolation_energy_grid.cpp:4600).
void orc_interpolate_hessian(const double* cell9, int rect, int nx, int ny, int nz, const double* grid, int n,
                             const double* pos, double* e, double* grad, double* hess)
{
  Box unit = makeBox(cell9, rect);
  const int np[3] = {nx, ny, nz};
  for (int q = 0; q < n; ++q) interpolateOne(unit, np, grid, pos + 3 * q, e + q, grad + 3 * q, hess + 9 * q);
}

// The grid test of System::createFrameworkInterpolationGrids (src/raspakit/system_interpolation_grids.cpp:240-840,
// tricubic branch) at n points given as fractional coordinates of the unit cell: the interpolated energy and
// gradient of the LJ grid (and of the unit-charge Ewald grid times probe_charge, when coul_grid is non-null) against
// calculateTricubicFractionalAtPosition over the supercell, its fractional gradient turned Cartesian with the
// transposed inverse cell (:302-305); weights exp(-beta E_LJ) with beta in inverse internal units and E_LJ in K
// (:365, :545); the Ewald part over the points with E_LJ (K) below the overlap criterion (:420).
// out[64] = {n, LJ stats[31], n_coulomb, Coulomb stats[31]}; stats = {sum w, <E> table, <E> full, relative error,
// <grad> table[3], <grad> full[3], relative error[3], <hessian> table[6], <hessian> full[6], relative error[6]},
// energies in K, gradients in K/A, Hessians (xx xy xz yy yz zz) in K/A^2.  values = 8 (tricubic; Hessian entries
// zero) or 27 (triquintic: calculateTriquinticFractionalAtPosition and interpolateHessian,
// system_interpolation_grids.cpp:328-441).
void orc_grid_check(const double* cell9, int rect, int nx, int ny, int nz, int natoms, const double* pos,
                    const int* type, const double* charge, int ntypes, const double* pair_eps, const double* pair_sigma,
                    const double* pair_shift, const int* pair_lj, int probe, double rc_vdw, double rc_supercell,
                    double alpha, double overlap, const double* lj_grid, const double* coul_grid, double probe_charge,
                    double temperature, int n, const double* frac, int values, double* out)
{
  Box unit = makeBox(cell9, rect);
  double w[3];
  perpendicularWidths(unit.cell, w);
  int nc[3];
  for (int a = 0; a < 3; ++a) nc[a] = static_cast<int>(std::ceil(2.0 * rc_supercell / w[a]));
  std::vector<double> spos;
  std::vector<int> stype;
  std::vector<double> sq;
  AtomsView av{natoms, pos, type, charge};
  makeSuperCell(unit.cell, av, nc, spos, stype, sq);
  Box super = unit;
  for (int col = 0; col < 3; ++col)
    for (int r = 0; r < 3; ++r) super.cell.c[col][r] = unit.cell.c[col][r] * nc[col];
  super.inv = inverse(super.cell);
  FF ff{ntypes, pair_eps, pair_sigma, pair_shift, pair_lj};
  const int np[3] = {nx, ny, nz};
  const int ns = static_cast<int>(stype.size());
  const double beta = 1.0 / (U.kB * temperature);
  const double K = U.energyToKelvin;
  struct Sums
  {
    double w = 0, t = 0, f = 0, d2 = 0, f2 = 0, gt[3] = {0, 0, 0}, gf[3] = {0, 0, 0}, gd2[3] = {0, 0, 0},
           gf2[3] = {0, 0, 0};
    double ht[6] = {}, hf[6] = {}, hd2[6] = {}, hf2[6] = {};
  } sl, sc;
  const bool quintic = values == 27;
  auto addH = [](Sums& S, double wt, const double* ht, const double* hf)
  {
    for (int a = 0; a < 6; ++a)
    {
      S.ht[a] += ht[a] * wt;
      S.hf[a] += hf[a] * wt;
      S.hd2[a] += (hf[a] - ht[a]) * (hf[a] - ht[a]) * wt;
      S.hf2[a] += hf[a] * hf[a] * wt;
    }
  };
  // interpolated and analytic values of one grid type at p: energy, Cartesian gradient, Hessian (6 components)
  auto eval = [&](int gridType, const double* grid, const double p[3], double scale, double& et, double gt[3],
                  double htab[6], double& ef, double gf[3], double hfull[6])
  {
    if (!quintic)
    {
      double v[8];
      interpolateOne(unit, np, grid, p, &et, gt);
      tricubicFractionalAtPosition(gridType, ff, rc_vdw, alpha, super, p, probe, unit, ns, spos.data(), stype.data(),
                                   sq.data(), v);
      ef = v[0];
      for (int c = 0; c < 3; ++c) gf[c] = unit.inv.c[c][0] * v[1] + unit.inv.c[c][1] * v[2] + unit.inv.c[c][2] * v[3];
      for (int a = 0; a < 6; ++a) htab[a] = hfull[a] = 0.0;
    }
    else
    {
      double v[27], h9[9];
      interpolateQuinticOne(unit, np, grid, p, &et, gt, h9);
      quinticFractionalAtPosition(gridType, ff, rc_vdw, alpha, super, p, probe, unit, ns, spos.data(), stype.data(),
                                  sq.data(), v);
      ef = v[0];
      for (int c = 0; c < 3; ++c) gf[c] = unit.inv.c[c][0] * v[1] + unit.inv.c[c][1] * v[2] + unit.inv.c[c][2] * v[3];
      // inverseCell^T H inverseCell with H the fractional second derivatives (values 4-9)
      const double Hs[3][3] = {{v[4], v[5], v[6]}, {v[5], v[7], v[8]}, {v[6], v[8], v[9]}};
      double Hc[3][3];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
        {
          double acc = 0.0;
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) acc += unit.inv.c[i][a] * Hs[a][b] * unit.inv.c[j][b];
          Hc[i][j] = acc;
        }
      const int ii[6] = {0, 0, 0, 1, 1, 2}, jj[6] = {0, 1, 2, 1, 2, 2};
      for (int a = 0; a < 6; ++a)
      {
        htab[a] = h9[3 * ii[a] + jj[a]] * scale;
        hfull[a] = Hc[ii[a]][jj[a]] * scale;
      }
    }
    et *= scale;
    ef *= scale;
    for (int c = 0; c < 3; ++c)
    {
      gt[c] *= scale;
      gf[c] *= scale;
    }
  };
  auto add = [](Sums& S, double wt, double et, double ef, const double* gt, const double* gf)
  {
    S.w += wt;
    S.t += et * wt;
    S.f += ef * wt;
    S.d2 += (ef - et) * (ef - et) * wt;
    S.f2 += ef * ef * wt;
    for (int a = 0; a < 3; ++a)
    {
      S.gt[a] += gt[a] * wt;
      S.gf[a] += gf[a] * wt;
      S.gd2[a] += (gf[a] - gt[a]) * (gf[a] - gt[a]) * wt;
      S.gf2[a] += gf[a] * gf[a] * wt;
    }
  };
  long long nq = 0;
  for (int i = 0; i < n; ++i)
  {
    double p[3];
    mulv(unit.cell, frac + 3 * i, p);
    double et, gt[3], ht[6], ef, gf[3], hf[6];
    eval(0, lj_grid, p, K, et, gt, ht, ef, gf, hf);
    const double wt = std::exp(-beta * ef);
    add(sl, wt, et, ef, gt, gf);
    addH(sl, wt, ht, hf);
    if (coul_grid && ef < overlap)
    {
      double ct, cgt[3], cht[6], cf, cgf[3], chf[6];
      eval(1, coul_grid, p, probe_charge * K, ct, cgt, cht, cf, cgf, chf);
      add(sc, wt, ct, cf, cgt, cgf);
      addH(sc, wt, cht, chf);
      ++nq;
    }
  }
  auto store = [](const Sums& S, double* o)
  {
    o[0] = S.w;
    o[1] = S.t / S.w;
    o[2] = S.f / S.w;
    o[3] = std::sqrt(S.d2 / S.f2);
    for (int a = 0; a < 3; ++a)
    {
      o[4 + a] = S.gt[a] / S.w;
      o[7 + a] = S.gf[a] / S.w;
      o[10 + a] = std::sqrt(S.gd2[a] / S.gf2[a]);
    }
    for (int a = 0; a < 6; ++a)
    {
      o[13 + a] = S.ht[a] / S.w;
      o[19 + a] = S.hf[a] / S.w;
      o[25 + a] = S.hf2[a] > 0.0 ? std::sqrt(S.hd2[a] / S.hf2[a]) : 0.0;
    }
  };
  for (int k = 0; k < 64; ++k) out[k] = 0.0;
  out[0] = n;
  store(sl, out + 1);
  out[32] = double(nq);
  if (nq > 0) store(sc, out + 33);
}

// The 64x64 tricubic coefficient matrix (row-major), for tests.
void orc_tricubic_matrix(double* out)
{
  for (int r = 0; r < 64; ++r)
    for (int m = 0; m < 64; ++m) out[64 * r + m] = TM.A[r][m];
}

// CBMC::computeFrameworkMoleculeEnergy, serial branch (src/raspakit/cbmc/interactions_framework_molecule.cpp:31-111)
// for one set of sites (a trial first bead or a trial chain set).  Sites whose type has a grid use it
// (grid_of_type[type] >= 0 selects the grid pointer), with the unit-charge Coulomb grid multiplied by the
// charge; other sites use the direct min-image sum in the system box.  Returns 1 on overlap (the reference
// returns std::nullopt), else 0 with energies[0] = VDW, energies[1] = Coulomb.
int orc_framework_energy(const double* box_cell9, int box_rect, int nfw, const double* fw_pos, const int* fw_type,
                         const double* fw_charge, const double* fw_scaling_vdw, const double* fw_scaling_coul,
                         int ntypes, const double* pair_eps, const double* pair_sigma, const double* pair_shift,
                         const int* pair_lj, double rc_vdw, double rc_coulomb, double alpha, int use_charge,
                         double overlap, int nsites, const double* site_pos, const int* site_type,
                         const double* site_charge, const double* site_scaling_vdw,
                         const double* site_scaling_coul, int skip, const double* grid_cell9, int grid_rect,
                         const int* grid_dims, const int* grid_of_type, const double* const* grids,
                         const double* coulomb_grid, double* energies)
{
  Box box = makeBox(box_cell9, box_rect);
  FF ff{ntypes, pair_eps, pair_sigma, pair_shift, pair_lj};
  const double rcv2 = rc_vdw * rc_vdw, rcc2 = rc_coulomb * rc_coulomb;
  double evdw = 0.0, ecoul = 0.0;
  for (int b = 0; b < nsites; ++b)
  {
    if (b == skip) continue;
    int tb = site_type[b];
    if (grid_of_type && grid_of_type[tb] >= 0)
    {
      Box gunit = makeBox(grid_cell9, grid_rect);
      double e;
      interpolateOne(gunit, grid_dims, grids[grid_of_type[tb]], site_pos + 3 * b, &e, nullptr);
      if (e > overlap) return 1;
      evdw += e;
      if (use_charge)
      {
        double c;
        interpolateOne(gunit, grid_dims, coulomb_grid, site_pos + 3 * b, &c, nullptr);
        ecoul += site_charge[b] * c;
      }
      continue;
    }
    for (int a = 0; a < nfw; ++a)
    {
      double dr[3] = {fw_pos[3 * a] - site_pos[3 * b], fw_pos[3 * a + 1] - site_pos[3 * b + 1],
                      fw_pos[3 * a + 2] - site_pos[3 * b + 2]};
      applyPBC(box, dr);
      double rr = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
      if (rr < rcv2)
      {
        double e = vdwEnergy(ff(fw_type[a], tb), fw_scaling_vdw[a] * site_scaling_vdw[b], rr);
        if (e > overlap) return 1;
        evdw += e;
      }
      if (use_charge && rr < rcc2)
      {
        double r = std::sqrt(rr);
        ecoul += coulombEnergy(alpha, fw_scaling_coul[a] * site_scaling_coul[b], r, fw_charge[a], site_charge[b]);
      }
    }
  }
  energies[0] = evdw;
  energies[1] = ecoul;
  return 0;
}

// CBMC::computeInterMolecularEnergy (src/raspakit/cbmc/interactions_intermolecular.cpp:22-95).
// 'skip_molecule' >= 0 excludes that background molecule (skipBackgroundMolecule).  Returns 1 on overlap.
int orc_intermolecular_energy(const double* box_cell9, int box_rect, int nmol_atoms, const double* m_pos,
                              const int* m_type, const double* m_charge, const double* m_scaling_vdw,
                              const double* m_scaling_coul, const int* m_molecule, int ntypes,
                              const double* pair_eps, const double* pair_sigma, const double* pair_shift,
                              const int* pair_lj, double rc_vdw, double rc_coulomb, double alpha, int use_charge,
                              double overlap, int nsites, const double* site_pos, const int* site_type,
                              const double* site_charge, const double* site_scaling_vdw,
                              const double* site_scaling_coul, const int* site_molecule, int skip,
                              int skip_molecule, double* energies)
{
  Box box = makeBox(box_cell9, box_rect);
  FF ff{ntypes, pair_eps, pair_sigma, pair_shift, pair_lj};
  const double rcv2 = rc_vdw * rc_vdw, rcc2 = rc_coulomb * rc_coulomb;
  double evdw = 0.0, ecoul = 0.0;
  for (int a = 0; a < nmol_atoms; ++a)
  {
    if (skip_molecule >= 0 && m_molecule[a] == skip_molecule) continue;
    for (int b = 0; b < nsites; ++b)
    {
      if (b == skip) continue;
      if (m_molecule[a] == site_molecule[b]) continue;
      double dr[3] = {m_pos[3 * a] - site_pos[3 * b], m_pos[3 * a + 1] - site_pos[3 * b + 1],
                      m_pos[3 * a + 2] - site_pos[3 * b + 2]};
      applyPBC(box, dr);
      double rr = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
      if (rr < rcv2)
      {
        double e = vdwEnergy(ff(m_type[a], site_type[b]), m_scaling_vdw[a] * site_scaling_vdw[b], rr);
        if (e > overlap) return 1;
        evdw += e;
      }
      if (use_charge && rr < rcc2)
      {
        double r = std::sqrt(rr);
        ecoul += coulombEnergy(alpha, m_scaling_coul[a] * site_scaling_coul[b], r, m_charge[a], site_charge[b]);
      }
    }
  }
  energies[0] = evdw;
  energies[1] = ecoul;
  return 0;
}

// Structure factor S_k = sum_i scaling_i q_i exp(i k.r_i) over the k-vectors in the reference's nvec order
// (kx = 0..kmax, ky = -kmax..kmax, kz = -kmax..kmax, skip k = 0 and the two cutoffs), with eik by recursion
// (src/raspakit/interactions/ewald_kvector.ixx:25-70, ewald.cpp:276-348 for the rigid framework part).
// Writes nk complex values (re, im) into s_out if non-null and returns nk.
static int ewaldLoop(const Box& box, double alpha, const int* kmax, long long recip_int_cut2, double recip_cut2,
                     int n, const double* pos, const double* q, double* s_out, double* temps)
{
  const int nkx = kmax[0], nky = kmax[1], nkz = kmax[2];
  std::vector<std::complex<double>> ex((nkx + 1) * std::size_t(n)), ey((nky + 1) * std::size_t(n)),
      ez((nkz + 1) * std::size_t(n));
  for (int i = 0; i < n; ++i)
  {
    double s[3];
    mulv(box.inv, pos + 3 * i, s);
    for (int a = 0; a < 3; ++a) s[a] *= 2.0 * M_PI;
    ex[i] = ey[i] = ez[i] = {1.0, 0.0};
    if (nkx >= 1) ex[i + n] = {std::cos(s[0]), std::sin(s[0])};
    if (nky >= 1) ey[i + n] = {std::cos(s[1]), std::sin(s[1])};
    if (nkz >= 1) ez[i + n] = {std::cos(s[2]), std::sin(s[2])};
  }
  for (int k = 2; k <= nkx; ++k)
    for (int i = 0; i < n; ++i) ex[i + k * n] = ex[i + (k - 1) * n] * ex[i + n];
  for (int k = 2; k <= nky; ++k)
    for (int i = 0; i < n; ++i) ey[i + k * n] = ey[i + (k - 1) * n] * ey[i + n];
  for (int k = 2; k <= nkz; ++k)
    for (int i = 0; i < n; ++i) ez[i + k * n] = ez[i + (k - 1) * n] * ez[i + n];
  // reciprocal vectors = rows of the inverse cell
  double ax[3] = {box.inv.c[0][0], box.inv.c[1][0], box.inv.c[2][0]};
  double ay[3] = {box.inv.c[0][1], box.inv.c[1][1], box.inv.c[2][1]};
  double az[3] = {box.inv.c[0][2], box.inv.c[1][2], box.inv.c[2][2]};
  const double prefactor = U.coulomb * (2.0 * M_PI / box.volume);
  std::vector<std::complex<double>> exy(n);
  int nvec = 0;
  for (int kx = 0; kx <= nkx; ++kx)
  {
    double factor = (kx == 0) ? prefactor : 2.0 * prefactor;
    for (int ky = -nky; ky <= nky; ++ky)
    {
      for (int i = 0; i < n; ++i)
      {
        std::complex<double> t = ey[i + n * std::abs(ky)];
        if (ky < 0) t = std::conj(t);
        exy[i] = ex[i + n * kx] * t;
      }
      for (int kz = -nkz; kz <= nkz; ++kz)
      {
        double kv[3];
        for (int a = 0; a < 3; ++a) kv[a] = 2.0 * M_PI * kx * ax[a] + 2.0 * M_PI * ky * ay[a] + 2.0 * M_PI * kz * az[a];
        double rksq = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        long long ksq = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
        if (ksq != 0 && ksq <= recip_int_cut2 && rksq < recip_cut2)
        {
          std::complex<double> ck(0.0, 0.0);
          for (int i = 0; i < n; ++i)
          {
            std::complex<double> t = ez[i + n * std::abs(kz)];
            if (kz < 0) t = std::conj(t);
            ck += q[i] * (exy[i] * t);
          }
          if (s_out)
          {
            s_out[2 * nvec] = ck.real();
            s_out[2 * nvec + 1] = ck.imag();
          }
          if (temps) temps[nvec] = factor * std::exp((-0.25 / (alpha * alpha)) * rksq) / rksq;
          ++nvec;
        }
      }
    }
  }
  return nvec;
}

// Number of k-vectors and the per-k prefactors f_k (incl. the kx>0 factor 2) for a box.
int orc_ewald_kvectors(const double* cell9, int rect, double alpha, const int* kmax, long long recip_int_cut2,
                       double recip_cut2, double* temps)
{
  Box box = makeBox(cell9, rect);
  double dummyp[3] = {0, 0, 0}, dummyq = 0.0;
  return ewaldLoop(box, alpha, kmax, recip_int_cut2, recip_cut2, 1, dummyp, &dummyq, nullptr, temps);
}

// Structure factors (scaled charges) of a set of atoms.
int orc_structure_factor(const double* cell9, int rect, double alpha, const int* kmax, long long recip_int_cut2,
                         double recip_cut2, int n, const double* pos, const double* q, double* s_out)
{
  Box box = makeBox(cell9, rect);
  return ewaldLoop(box, alpha, kmax, recip_int_cut2, recip_cut2, n, pos, q, s_out, nullptr);
}

// Interactions::computeEwaldFourierEnergySingleIon (src/raspakit/interactions/ewald.cpp:210-272).
double orc_ewald_single_ion(const double* cell9, int rect, double alpha, const int* kmax, long long recip_int_cut2,
                            double recip_cut2)
{
  Box box = makeBox(cell9, rect);
  std::vector<double> temps(4 * (kmax[0] + 1) * (kmax[1] + 1) * (kmax[2] + 1) + 8);
  double p[3] = {0, 0, 0}, q = 1.0;
  std::vector<double> s(2 * temps.size());
  int nk = ewaldLoop(box, alpha, kmax, recip_int_cut2, recip_cut2, 1, p, &q, s.data(), temps.data());
  // temps already carry prefactor = C*2pi/V ; energy_sum in the reference is without it
  double sum = 0.0;
  for (int k = 0; k < nk; ++k) sum += temps[k] * (s[2 * k] * s[2 * k] + s[2 * k + 1] * s[2 * k + 1]);
  return -(sum - U.coulomb * alpha / std::sqrt(M_PI));
}

// Interactions::computeEwaldFourierEnergy (src/raspakit/interactions/ewald.cpp:352-563) without
// fractional groups: Fourier (total minus rigid framework), self, intramolecular exclusion and the net-charge
// term.  Molecules are consecutive runs of equal molecule id in the adsorbate arrays.
// out[0] = ewald_fourier, out[1] = ewald_self, out[2] = ewald_exclusion; writes the stored structure factor
// (framework + adsorbates) into stored_out (2*nk doubles) when non-null.
int orc_ewald_total(const double* cell9, int rect, double alpha, const int* kmax, long long recip_int_cut2,
                    double recip_cut2, int nfw, const double* fw_pos, const double* fw_q, int nads,
                    const double* a_pos, const double* a_q, const double* a_scaling, const int* a_molecule,
                    double net_charge_framework, double* out, double* stored_out)
{
  Box box = makeBox(cell9, rect);
  std::size_t cap = 4 * std::size_t(kmax[0] + 1) * (kmax[1] + 1) * (kmax[2] + 1) + 8;
  std::vector<double> sf(2 * cap), sa(2 * cap), temps(cap);
  std::vector<double> qa(nads);
  for (int i = 0; i < nads; ++i) qa[i] = a_scaling[i] * a_q[i];
  int nk = ewaldLoop(box, alpha, kmax, recip_int_cut2, recip_cut2, nfw, fw_pos, fw_q, sf.data(), temps.data());
  ewaldLoop(box, alpha, kmax, recip_int_cut2, recip_cut2, nads, a_pos, qa.data(), sa.data(), nullptr);
  double ef = 0.0, single = 0.0;
  for (int k = 0; k < nk; ++k)
  {
    single += temps[k];
    double tr = sf[2 * k] + sa[2 * k], ti = sf[2 * k + 1] + sa[2 * k + 1];
    double rig = temps[k] * (sf[2 * k] * sf[2 * k] + sf[2 * k + 1] * sf[2 * k + 1]);
    ef += temps[k] * (tr * tr + ti * ti) - rig;
    if (stored_out)
    {
      stored_out[2 * k] = tr;
      stored_out[2 * k + 1] = ti;
    }
  }
  double eself = 0.0;
  const double pself = U.coulomb * alpha / std::sqrt(M_PI);
  for (int i = 0; i < nads; ++i) eself -= pself * qa[i] * qa[i];
  double eexcl = 0.0;
  for (int i = 0; i < nads; ++i)
    for (int j = i + 1; j < nads && a_molecule[j] == a_molecule[i]; ++j)
    {
      double dr[3] = {a_pos[3 * i] - a_pos[3 * j], a_pos[3 * i + 1] - a_pos[3 * j + 1], a_pos[3 * i + 2] - a_pos[3 * j + 2]};
      applyPBC(box, dr);
      double r = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      double st = a_scaling[i] * a_scaling[j];
      double s = r + 0.01 * (1.0 - st);
      double pot = std::erf(alpha * s) / s;
      eexcl -= st * U.coulomb * a_q[i] * a_q[j] * pot;
    }
  double netAds = 0.0;
  for (int i = 0; i < nads; ++i) netAds += qa[i];
  double uIon = -(single - U.coulomb * alpha / std::sqrt(M_PI));
  ef += uIon * (2.0 * net_charge_framework + netAds) * netAds;
  out[0] = ef;
  out[1] = eself;
  out[2] = eexcl;
  return nk;
}

// Interactions::energyDifferenceEwaldFourier (src/raspakit/interactions/ewald.cpp:1043-1244) without
// fractional groups.  'stored' holds S_k (2*nk doubles); trial_out receives S_k + dS_k.  Old and new atoms
// carry their own charges and Coulomb scalings.  out[0] = fourier, out[1] = self, out[2] = exclusion.
int orc_ewald_delta(const double* cell9, int rect, double alpha, const int* kmax, long long recip_int_cut2,
                    double recip_cut2, const double* stored, int nnew, const double* n_pos, const double* n_q,
                    const double* n_scaling, const int* n_molecule, int nold, const double* o_pos,
                    const double* o_q, const double* o_scaling, const int* o_molecule, double net_charge,
                    double* out, double* trial_out)
{
  Box box = makeBox(cell9, rect);
  std::size_t cap = 4 * std::size_t(kmax[0] + 1) * (kmax[1] + 1) * (kmax[2] + 1) + 8;
  std::vector<double> sn(2 * cap, 0.0), so(2 * cap, 0.0), temps(cap);
  std::vector<double> qn(nnew), qo(nold);
  for (int i = 0; i < nnew; ++i) qn[i] = n_scaling[i] * n_q[i];
  for (int i = 0; i < nold; ++i) qo[i] = o_scaling[i] * o_q[i];
  double p0[3] = {0, 0, 0}, q0 = 0.0;
  int nk = ewaldLoop(box, alpha, kmax, recip_int_cut2, recip_cut2, 1, p0, &q0, nullptr, temps.data());
  if (nnew) ewaldLoop(box, alpha, kmax, recip_int_cut2, recip_cut2, nnew, n_pos, qn.data(), sn.data(), nullptr);
  if (nold) ewaldLoop(box, alpha, kmax, recip_int_cut2, recip_cut2, nold, o_pos, qo.data(), so.data(), nullptr);
  double ef = 0.0, single = 0.0;
  for (int k = 0; k < nk; ++k)
  {
    single += temps[k];
    double tr = stored[2 * k] + sn[2 * k] - so[2 * k];
    double ti = stored[2 * k + 1] + sn[2 * k + 1] - so[2 * k + 1];
    ef += temps[k] * (tr * tr + ti * ti);
    ef -= temps[k] * (stored[2 * k] * stored[2 * k] + stored[2 * k + 1] * stored[2 * k + 1]);
    if (trial_out)
    {
      trial_out[2 * k] = tr;
      trial_out[2 * k + 1] = ti;
    }
  }
  auto excl = [&](int n, const double* p, const double* q, const double* sc, const int* mol) -> double
  {
    double e = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j)
      {
        if (mol[i] != mol[j]) continue;
        double dr[3] = {p[3 * i] - p[3 * j], p[3 * i + 1] - p[3 * j + 1], p[3 * i + 2] - p[3 * j + 2]};
        applyPBC(box, dr);
        double r = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
        double st = sc[i] * sc[j];
        double s = r + 0.01 * (1.0 - st);
        e += st * U.coulomb * q[i] * q[j] * (std::erf(alpha * s) / s);
      }
    return e;
  };
  double eexcl = excl(nold, o_pos, o_q, o_scaling, o_molecule) - excl(nnew, n_pos, n_q, n_scaling, n_molecule);
  const double pself = U.coulomb * alpha / std::sqrt(M_PI);
  double eself = 0.0;
  for (int i = 0; i < nold; ++i) eself += pself * qo[i] * qo[i];
  for (int i = 0; i < nnew; ++i) eself -= pself * qn[i] * qn[i];
  double dq = 0.0;
  for (int i = 0; i < nold; ++i) dq -= qo[i];
  for (int i = 0; i < nnew; ++i) dq += qn[i];
  double uIon = -(single - U.coulomb * alpha / std::sqrt(M_PI));
  double netNew = net_charge + dq;
  ef += uIon * (netNew * netNew - net_charge * net_charge);
  out[0] = ef;
  out[1] = eself;
  out[2] = eexcl;
  return nk;
}

// CBMC::selectTrialPosition (src/raspakit/cbmc/util.cpp:15-48) given the uniform it draws.
int orc_select_trial(int n, const double* lbf, double uniform)
{
  double largest = lbf[0];
  for (int i = 1; i < n; ++i)
    if (lbf[i] > largest) largest = lbf[i];
  std::vector<double> w(n);
  double sum = 0.0;
  for (int i = 0; i < n; ++i)
  {
    w[i] = std::exp(lbf[i] - largest);
    sum += w[i];
  }
  int sel = 0;
  double cumw = w[0];
  double ws = uniform * sum;
  while (sel + 1 < n && cumw < ws) cumw += w[++sel];
  return sel;
}

}  // extern "C"

#include "oracle_mc.inc"

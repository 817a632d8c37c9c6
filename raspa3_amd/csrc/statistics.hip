// Statistics the reference reports for a Monte Carlo run, restated as host code behind the C ABI: block averages
// with their Student-t confidence intervals, the enthalpy of adsorption of every swappable component from the
// fluctuation matrix, and the Widom chemical potential and fugacity.  No device work.
//
//   block error            src/raspakit/averages.ixx:28-46 (t-table, 5 blocks, 95 %), :144-192
//   block average          src/raspakit/properties/property_block_average.ixx:54-150
//   enthalpy of adsorption src/raspakit/properties/enthalpy.ixx:101-200
//   Widom                  src/raspakit/properties/widom.ixx:155-230
#include <algorithm>
#include <cmath>
#include <vector>

#include "rg_internal.hpp"

namespace
{

// Two-sided Student-t critical values at 95 % for 0..20 degrees of freedom (the 95 % column of the reference's table
// of deviates; entry 0 is unused).
constexpr double kStudentT95[21] = {0.0,   12.71, 4.303, 3.182, 2.776, 2.571, 2.447, 2.365, 2.306, 2.262, 2.228,
                                    2.201, 2.179, 2.160, 2.145, 2.131, 2.120, 2.110, 2.101, 2.093, 2.086};

double blockError(const std::vector<double>& blocks, double average)
{
  const size_t n = blocks.size();
  if (n < 3) return 0.0;
  double ss = 0.0;
  for (double b : blocks) ss += (b - average) * (b - average);
  const double sd = std::sqrt(ss / double(n - 1));
  return kStudentT95[n - 1] * sd / std::sqrt(double(n));
}

// Blocks used for an error estimate: weight above half the first block's (the reference's "at least half filled").
std::vector<int> filledBlocks(int n, const double* weight)
{
  std::vector<int> use;
  const double ref = std::max(1.0, n > 0 ? weight[0] : 1.0);
  for (int b = 0; b < n; ++b)
    if (weight[b] / ref > 0.5) use.push_back(b);
  return use;
}

// Enthalpy of adsorption of every component from one set of averaged terms: the symmetric fluctuation matrix
// M_ij = <N_i N_j> - <N_i><N_j> is inverted and v_i = sum_j (<U N_j> - <U><N_j>) (M^-1)_ji - kT.  Gauss-Jordan with
// partial pivoting; a singular matrix gives NaN.
void enthalpyOfTerms(int nc, double u, const double* n, const double* un, const double* nn, double kT, double* out)
{
  std::vector<double> m(size_t(nc) * nc), inv(size_t(nc) * nc, 0.0);
  for (int i = 0; i < nc; ++i)
  {
    for (int j = 0; j < nc; ++j) m[i * nc + j] = nn[i * nc + j] - n[i] * n[j];
    inv[i * nc + i] = 1.0;
  }
  bool singular = false;
  for (int c = 0; c < nc && !singular; ++c)
  {
    int piv = c;
    for (int r = c + 1; r < nc; ++r)
      if (std::fabs(m[r * nc + c]) > std::fabs(m[piv * nc + c])) piv = r;
    if (!(std::fabs(m[piv * nc + c]) > 0.0))
    {
      singular = true;
      break;
    }
    for (int k = 0; k < nc; ++k)
    {
      std::swap(m[c * nc + k], m[piv * nc + k]);
      std::swap(inv[c * nc + k], inv[piv * nc + k]);
    }
    const double d = 1.0 / m[c * nc + c];
    for (int k = 0; k < nc; ++k)
    {
      m[c * nc + k] *= d;
      inv[c * nc + k] *= d;
    }
    for (int r = 0; r < nc; ++r)
    {
      if (r == c) continue;
      const double f = m[r * nc + c];
      for (int k = 0; k < nc; ++k)
      {
        m[r * nc + k] -= f * m[c * nc + k];
        inv[r * nc + k] -= f * inv[c * nc + k];
      }
    }
  }
  for (int i = 0; i < nc; ++i)
  {
    if (singular)
    {
      out[i] = std::nan("");
      continue;
    }
    double v = 0.0;
    for (int j = 0; j < nc; ++j) v += (un[j] - u * n[j]) * inv[j * nc + i];
    out[i] = v - kT;
  }
}

}  // namespace

extern "C"
{

rg_status rg_block_error(int n, const double* block_averages, double average, double* ci)
{
  if (n < 0 || n > 21 || (n > 0 && !block_averages) || !ci) return RG_ERR_INVALID;
  *ci = blockError(std::vector<double>(block_averages, block_averages + n), average);
  return RG_OK;
}

rg_status rg_block_average(int n_blocks, const double* sums, const double* weights, double* mean, double* ci)
{
  if (n_blocks < 1 || n_blocks > 21 || !sums || !weights || !mean || !ci) return RG_ERR_INVALID;
  double s = 0.0, w = 0.0;
  for (int b = 0; b < n_blocks; ++b)
  {
    s += sums[b];
    w += weights[b];
  }
  *mean = s / std::max(1.0, w);
  std::vector<double> blocks;
  for (int b : filledBlocks(n_blocks, weights)) blocks.push_back(sums[b] / std::max(1.0, weights[b]));
  *ci = blockError(blocks, *mean);
  return RG_OK;
}

rg_status rg_enthalpy_of_adsorption(int n_blocks, int n_components, const double* weight, const double* sum_u,
                                    const double* sum_n, const double* sum_un, const double* sum_nn,
                                    double temperature, double* mean, double* ci)
{
  if (n_blocks < 1 || n_blocks > 21 || n_components < 1 || n_components > 16 || !weight || !sum_u || !sum_n ||
      !sum_un || !sum_nn || !mean || !ci || !(temperature > 0.0))
    return RG_ERR_INVALID;
  const int nc = n_components;
  const double kT = temperature * rg::units().kelvinToEnergy;
  const std::vector<int> use = filledBlocks(n_blocks, weight);
  std::vector<std::vector<double>> perBlock;
  std::vector<double> n(nc), un(nc), nn(size_t(nc) * nc), v(nc);
  for (int b : use)
  {
    const double w = std::max(1.0, weight[b]);
    for (int i = 0; i < nc; ++i)
    {
      n[i] = sum_n[b * nc + i] / w;
      un[i] = sum_un[b * nc + i] / w;
      for (int j = 0; j < nc; ++j) nn[i * nc + j] = sum_nn[(size_t(b) * nc + i) * nc + j] / w;
    }
    enthalpyOfTerms(nc, sum_u[b] / w, n.data(), un.data(), nn.data(), kT, v.data());
    perBlock.push_back(v);
  }
  // the overall value is the mean of the block composites (the fluctuation formula is not linear in the averages)
  for (int i = 0; i < nc; ++i)
  {
    double s = 0.0;
    for (const auto& pb : perBlock) s += pb[i];
    mean[i] = s / double(std::max<size_t>(1, perBlock.size()));
    std::vector<double> col;
    for (const auto& pb : perBlock) col.push_back(pb[i]);
    ci[i] = blockError(col, mean[i]);
  }
  return RG_OK;
}

rg_status rg_widom_statistics(int n_blocks, const double* weight, const double* sum_w, const double* sum_density,
                              double temperature, rg_widom_stats_t* out)
{
  if (n_blocks < 1 || n_blocks > 21 || !weight || !sum_w || !sum_density || !out || !(temperature > 0.0))
    return RG_ERR_INVALID;
  const double kT = temperature * rg::units().kelvinToEnergy;
  struct Mu
  {
    double ex, id, tot, fug;
  };
  auto transform = [kT](double w, double rho) {
    Mu m;
    m.ex = -kT * std::log(w);
    m.id = kT * std::log(rho);
    m.tot = m.ex + m.id;
    m.fug = std::exp(m.tot / kT) * kT;
    return m;
  };
  double sw = 0.0, sr = 0.0, wt = 0.0;
  for (int b = 0; b < n_blocks; ++b)
  {
    sw += sum_w[b];
    sr += sum_density[b];
    wt += weight[b];
  }
  const double wMean = sw / std::max(1.0, wt), rMean = sr / std::max(1.0, wt);
  const Mu mean = transform(wMean, rMean);
  std::vector<double> bw, bex, bid, btot, bfug;
  for (int b : filledBlocks(n_blocks, weight))
  {
    const double w = std::max(1.0, weight[b]);
    const Mu m = transform(sum_w[b] / w, sum_density[b] / w);
    bw.push_back(sum_w[b] / w);
    bex.push_back(m.ex);
    bid.push_back(m.id);
    btot.push_back(m.tot);
    bfug.push_back(m.fug);
  }
  out->rosenbluth_weight = wMean;
  out->rosenbluth_weight_ci = blockError(bw, wMean);
  out->mu_excess = mean.ex;
  out->mu_excess_ci = blockError(bex, mean.ex);
  out->mu_ideal_gas = mean.id;
  out->mu_ideal_gas_ci = blockError(bid, mean.id);
  out->mu_total = mean.tot;
  out->mu_total_ci = blockError(btot, mean.tot);
  out->fugacity = mean.fug;
  out->fugacity_ci = blockError(bfug, mean.fug);
  return RG_OK;
}

}  // extern "C"

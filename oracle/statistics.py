"""Statistics of a Monte Carlo run as the reference computes them, restated in numpy (test infrastructure only: the
checker of the library's rg_block_error, rg_block_average, rg_enthalpy_of_adsorption and rg_widom_statistics).

Follows:
  * Student-t deviates, 5 blocks, 95 %: src/raspakit/averages.ixx:28-46
  * blockErrorEstimate and the "at least half filled" block selection: src/raspakit/averages.ixx:144-192
  * BlockAverage (direct: pooled weighted mean; composite: mean of the block composites):
    src/raspakit/properties/property_block_average.ixx:54-150
  * EnthalpyOfAdsorptionTerms::compositeProperty: src/raspakit/properties/enthalpy.ixx:177-200
  * PropertyWidom chemical potential and fugacity transforms: src/raspakit/properties/widom.ixx:176-230

Pinned by (tests/test_statistics.py): the t deviates against scipy's Student-t quantiles; the interval's 95 %
coverage on Gaussian blocks of known mean; the enthalpy formula against exact regression coefficients of a linear
energy model (U = a N1 + b N2 + c gives a - kT and b - kT) and its one-component reduction to
(<UN> - <U><N>) / (<N^2> - <N>^2) - kT; mu_excess = 0 for W = 1 and the ideal-gas fugacity rho kT.
"""
import numpy as np

# 95 % column of the reference's deviates for 0..20 degrees of freedom (entry 0 unused)
T95 = np.array([0.0, 12.71, 4.303, 3.182, 2.776, 2.571, 2.447, 2.365, 2.306, 2.262, 2.228, 2.201, 2.179, 2.160,
                2.145, 2.131, 2.120, 2.110, 2.101, 2.093, 2.086])
KELVIN_TO_ENERGY = 0.8314457542157568  # Boltzmann's constant in internal energy units per K (src/raspakit/units.cpp)


def block_error(blocks, average):
    """averages.ixx:144-169: t_(n-1) * s / sqrt(n); 0 for fewer than 3 blocks."""
    b = np.asarray(blocks, float)
    n = len(b)
    if n < 3:
        return 0.0 * average
    sd = np.sqrt(((b - average) ** 2).sum(0) / (n - 1))
    return T95[n - 1] * sd / np.sqrt(n)


def filled(weights):
    """averages.ixx:180-192: blocks whose weight exceeds half the first block's (reference weight at least 1)."""
    w = np.asarray(weights, float)
    ref = max(1.0, w[0]) if len(w) else 1.0
    return [b for b in range(len(w)) if w[b] / ref > 0.5]


def block_average(sums, weights):
    """property_block_average.ixx:108-150 for a direct property: pooled mean, interval over the filled blocks."""
    s, w = np.asarray(sums, float), np.asarray(weights, float)
    mean = s.sum() / max(1.0, w.sum())
    blocks = [s[b] / max(1.0, w[b]) for b in filled(w)]
    return mean, block_error(blocks, mean)


def enthalpy_terms(u, n, un, nn, kT):
    """enthalpy.ixx:177-200 for one set of averaged terms."""
    n, un, nn = np.asarray(n, float), np.asarray(un, float), np.asarray(nn, float)
    m = nn - np.outer(n, n)
    minv = np.linalg.inv(m)
    return (un - u * n) @ minv - kT


def enthalpy_of_adsorption(weight, sum_u, sum_n, sum_un, sum_nn, temperature):
    """Composite block average of the enthalpy terms: per filled block the composite, then mean and interval over
    those blocks (property_block_average.ixx:108-126)."""
    kT = temperature * KELVIN_TO_ENERGY
    w = np.asarray(weight, float)
    per = []
    for b in filled(w):
        wb = max(1.0, w[b])
        per.append(enthalpy_terms(sum_u[b] / wb, np.asarray(sum_n[b]) / wb, np.asarray(sum_un[b]) / wb,
                                  np.asarray(sum_nn[b]) / wb, kT))
    per = np.array(per)
    mean = per.mean(0)
    return mean, block_error(per, mean)


def widom_statistics(weight, sum_w, sum_density, temperature):
    """widom.ixx:176-230: transforms of the pooled averages; intervals from the transforms of the block averages."""
    kT = temperature * KELVIN_TO_ENERGY
    w = np.asarray(weight, float)
    sw, sr = np.asarray(sum_w, float), np.asarray(sum_density, float)

    def transform(wr, rho):
        ex = -kT * np.log(wr)
        idg = kT * np.log(rho)
        return dict(rosenbluth_weight=wr, mu_excess=ex, mu_ideal_gas=idg, mu_total=ex + idg,
                    fugacity=np.exp((ex + idg) / kT) * kT)

    mean = transform(sw.sum() / max(1.0, w.sum()), sr.sum() / max(1.0, w.sum()))
    blocks = [transform(sw[b] / max(1.0, w[b]), sr[b] / max(1.0, w[b])) for b in filled(w)]
    out = {}
    for k, v in mean.items():
        out[k] = v
        out[k + "_ci"] = block_error([bl[k] for bl in blocks], v)
    return out

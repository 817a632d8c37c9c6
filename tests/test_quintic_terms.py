"""Pins the generated triquintic moment code (raspa3_amd/csrc/quintic_terms.hpp) independently of the partition
enumeration that generated it: the header is compiled for the host with g++ (the HIP qualifiers defined away),
fed one pair's radial factors f_0..f_6, offsets u and a metric G, and its 27 values are compared with the Taylor
coefficients of F(q0 + w), w = u.d + d^T G d / 2, expanded with sympy (F^(k)(q0) = f_k): value (a, b, c) is
a! b! c! times the coefficient of d0^a d1^b d2^c.  Both the triclinic form (general G) and the rectangular form
(G = identity) are checked.
"""
import math
import os
import shutil
import subprocess

import pytest
import sympy as sp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "raspa3_amd", "csrc")
ORDERS = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (1, 0, 1), (0, 2, 0), (0, 1, 1),
          (0, 0, 2), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2), (2, 2, 0),
          (2, 0, 2), (0, 2, 2), (2, 1, 1), (1, 2, 1), (1, 1, 2), (2, 2, 1), (2, 1, 2), (1, 2, 2), (2, 2, 2)]

F = [1.3, -0.7, 0.45, -0.21, 0.13, -0.077, 0.031]
U = [0.8, -1.1, 0.6]
G6 = [2.0, 3.0, 1.5, 0.4, -0.3, 0.25]  # G00 G11 G22 G01 G02 G12

HARNESS = r"""
#include <cmath>
#include <cstdio>
#include "quintic_terms.hpp"
int main()
{
  const double f[7] = {%s};
  const double G[6] = {%s};
  double S[rgq::kMomFull] = {0}, v[27];
  rgq::accumulate<true>(S, f, %s);
  rgq::combine<true>(S, G, v);
  for (int o = 0; o < 27; ++o) std::printf("%%.17g\n", v[o]);
  double S2[rgq::kMomFull] = {0};
  rgq::accumulate<false>(S2, f, %s);
  rgq::combine<false>(S2, G, v);
  for (int o = 0; o < 27; ++o) std::printf("%%.17g\n", v[o]);
}
"""


def taylor_values(G):
    d = sp.symbols("d0:3")
    q = sum(sp.Rational(str(U[a])) * d[a] for a in range(3))
    q += sp.Rational(1, 2) * sum(sp.Rational(str(G[a][b])) * d[a] * d[b] for a in range(3) for b in range(3))
    series = sp.expand(sum(sp.Rational(str(F[k])) * q ** k / math.factorial(k) for k in range(7)))
    poly = sp.Poly(series, *d)
    return [float(poly.coeff_monomial(d[0] ** a * d[1] ** b * d[2] ** c)) * math.factorial(a) * math.factorial(b)
            * math.factorial(c) for a, b, c in ORDERS]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_generated_moment_code_equals_taylor_coefficients(tmp_path):
    stub = tmp_path / "hip"
    stub.mkdir()
    (stub / "hip_runtime.h").write_text("")
    src = tmp_path / "h.cpp"
    fl = lambda xs: ", ".join(repr(x) for x in xs)
    src.write_text(HARNESS % (fl(F), fl(G6), fl(U), fl(U)))
    exe = tmp_path / "h"
    subprocess.run(["g++", "-O1", "-std=c++17", "-D__device__=", "-D__forceinline__=inline", f"-I{tmp_path}",
                    f"-I{CSRC}", str(src), "-o", str(exe)], check=True)
    out = [float(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    Gm = [[G6[0], G6[3], G6[4]], [G6[3], G6[1], G6[5]], [G6[4], G6[5], G6[2]]]
    ident = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    for got, G in ((out[:27], Gm), (out[27:], ident)):
        ref = taylor_values(G)
        for o, (a, b) in enumerate(zip(got, ref)):
            assert abs(a - b) <= 1e-13 * max(1.0, abs(b)), (o, ORDERS[o], a, b)

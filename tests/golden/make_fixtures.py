"""Writes the golden fixtures used by the oracle tests.

itq29.json: the ITQ-29 (LTA) unit cell of the reference's test fixture Framework::makeITQ29
(src/raspakit/framework.cpp:1561-1590): cubic a = 11.8671 A, space group Hall 517 (Pm-3m), asymmetric unit
  Si (0.3683, 0.1847, 0)      q = +2.05
  O  (0.5, 0.2179, 0)         q = -1.025
  O  (0.2939, 0.2939, 0)      q = -1.025
  O  (0.3429, 0.1098, 0.1098) q = -1.025
expanded here with the 48 operations of Pm-3m (all signed permutations of x, y, z; no translations), wrapped into
[0, 1) and de-duplicated (tolerance 1e-4 in fractional units) -> 24 Si + 48 O.  The reference expands the same
asymmetric unit with its own symmetry code (src/raspakit/cif_reader.cpp, CIFReader::expandDefinedAtomsToUnitCell);
only the order of the atoms can differ, which does not change any energy beyond rounding.

Run: python tests/golden/make_fixtures.py
"""
import itertools
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def pm3m_expand(asym, tol=1e-4):
    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            ops.append((perm, signs))
    out = []
    for (f, name, q) in asym:
        for perm, signs in ops:
            g = np.array([signs[a] * f[perm[a]] for a in range(3)])
            g = g - np.floor(g)
            g[np.abs(g - 1.0) < 1e-12] = 0.0
            dup = False
            for (h, _, _) in out:
                d = g - h
                d -= np.rint(d)
                if np.all(np.abs(d) < tol):
                    dup = True
                    break
            if not dup:
                out.append((g, name, q))
    return out


def main():
    asym = [((0.3683, 0.1847, 0.0), "Si", 2.05), ((0.5, 0.2179, 0.0), "O", -1.025),
            ((0.2939, 0.2939, 0.0), "O", -1.025), ((0.3429, 0.1098, 0.1098), "O", -1.025)]
    atoms = pm3m_expand(asym)
    data = dict(source="src/raspakit/framework.cpp:1561-1590 (Framework::makeITQ29), expanded with Pm-3m",
                a=11.8671, fractional=[list(map(float, a[0])) for a in atoms], element=[a[1] for a in atoms],
                charge=[a[2] for a in atoms])
    with open(os.path.join(HERE, "itq29.json"), "w") as f:
        json.dump(data, f, indent=1)
    print(len(atoms), "atoms:", sum(1 for a in atoms if a[1] == "Si"), "Si")


if __name__ == "__main__":
    main()

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

mfi.json: the MFI (silicalite-1) unit cell of Framework::makeMFI (src/raspakit/framework.cpp:1591-1689):
orthorhombic 20.022 x 19.899 x 13.383 A, space group Hall 292 (Pnma), 12 Si + 26 O in the asymmetric unit,
expanded with the 8 operations of Pnma -> 96 Si + 192 O (each Si with 4 O within 1.8 A, each O with 2 Si).

Both files are written here and into raspa3_amd/data/ (the workload generator reads them from there).

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


# asymmetric unit of Framework::makeMFI (fractional x, y, z; element; charge)
MFI_ASYM = [
    ((0.42238, 0.0565, -0.33598), "Si", 2.05),
    ((0.30716, 0.02772, -0.1893), "Si", 2.05),
    ((0.27911, 0.06127, 0.0312), "Si", 2.05),
    ((0.12215, 0.06298, 0.0267), "Si", 2.05),
    ((0.07128, 0.02722, -0.18551), "Si", 2.05),
    ((0.18641, 0.05896, -0.32818), "Si", 2.05),
    ((0.42265, -0.1725, -0.32718), "Si", 2.05),
    ((0.30778, -0.13016, -0.18548), "Si", 2.05),
    ((0.27554, -0.17279, 0.03109), "Si", 2.05),
    ((0.12058, -0.1731, 0.02979), "Si", 2.05),
    ((0.07044, -0.13037, -0.182), "Si", 2.05),
    ((0.18706, -0.17327, -0.31933), "Si", 2.05),
    ((0.3726, 0.0534, -0.2442), "O", -1.025),
    ((0.3084, 0.0587, -0.0789), "O", -1.025),
    ((0.2007, 0.0592, 0.0289), "O", -1.025),
    ((0.0969, 0.0611, -0.0856), "O", -1.025),
    ((0.1149, 0.0541, -0.2763), "O", -1.025),
    ((0.2435, 0.0553, -0.246), "O", -1.025),
    ((0.3742, -0.1561, -0.2372), "O", -1.025),
    ((0.3085, -0.1552, -0.0728), "O", -1.025),
    ((0.198, -0.1554, 0.0288), "O", -1.025),
    ((0.091, -0.1614, -0.0777), "O", -1.025),
    ((0.1169, -0.1578, -0.2694), "O", -1.025),
    ((0.2448, -0.1594, -0.2422), "O", -1.025),
    ((0.3047, -0.051, -0.1866), "O", -1.025),
    ((0.0768, -0.0519, -0.1769), "O", -1.025),
    ((0.4161, 0.1276, -0.3896), "O", -1.025),
    ((0.4086, -0.0017, -0.4136), "O", -1.025),
    ((0.402, -0.1314, -0.4239), "O", -1.025),
    ((0.1886, 0.1298, -0.3836), "O", -1.025),
    ((0.194, 0.0007, -0.4082), "O", -1.025),
    ((0.1951, -0.1291, -0.419), "O", -1.025),
    ((-0.0037, 0.0502, -0.208), "O", -1.025),
    ((-0.004, -0.1528, -0.2078), "O", -1.025),
    ((0.4192, -0.25, -0.354), "O", -1.025),
    ((0.1884, -0.25, -0.3538), "O", -1.025),
    ((0.2883, -0.25, 0.0579), "O", -1.025),
    ((0.1085, -0.25, 0.0611), "O", -1.025)
]


def pnma_ops(f):
    x, y, z = f
    return [(x, y, z), (-x + 0.5, -y, z + 0.5), (-x, y + 0.5, -z), (x + 0.5, -y + 0.5, -z + 0.5), (-x, -y, -z),
            (x + 0.5, y, -z + 0.5), (x, -y + 0.5, z), (-x + 0.5, y + 0.5, z + 0.5)]


def expand(asym, ops, tol=1e-4):
    out = []
    for (f, name, q) in asym:
        for g in ops(f):
            g = np.array(g, float)
            g = g - np.floor(g)
            g[np.abs(g - 1.0) < 1e-12] = 0.0
            if any(np.all(np.abs((g - h) - np.rint(g - h)) < tol) for (h, _, _) in out):
                continue
            out.append((g, name, q))
    return out


def write(name, data):
    for d in (HERE, os.path.join(os.path.dirname(os.path.dirname(HERE)), "raspa3_amd", "data")):
        with open(os.path.join(d, name), "w") as f:
            json.dump(data, f, indent=1)


def main():
    asym = [((0.3683, 0.1847, 0.0), "Si", 2.05), ((0.5, 0.2179, 0.0), "O", -1.025),
            ((0.2939, 0.2939, 0.0), "O", -1.025), ((0.3429, 0.1098, 0.1098), "O", -1.025)]
    atoms = pm3m_expand(asym)
    data = dict(source="src/raspakit/framework.cpp:1561-1590 (Framework::makeITQ29), expanded with Pm-3m",
                a=11.8671, fractional=[list(map(float, a[0])) for a in atoms], element=[a[1] for a in atoms],
                charge=[a[2] for a in atoms])
    write("itq29.json", data)
    print(len(atoms), "atoms:", sum(1 for a in atoms if a[1] == "Si"), "Si")
    mfi = expand(MFI_ASYM, pnma_ops)
    write("mfi.json", dict(source="src/raspakit/framework.cpp:1591-1689 (Framework::makeMFI), expanded with Pnma",
                           a=20.022, b=19.899, c=13.383, fractional=[list(map(float, a[0])) for a in mfi],
                           element=[a[1] for a in mfi], charge=[a[2] for a in mfi]))
    print(len(mfi), "atoms:", sum(1 for a in mfi if a[1] == "Si"), "Si")


if __name__ == "__main__":
    main()

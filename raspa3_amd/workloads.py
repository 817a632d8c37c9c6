"""Synthetic, seeded workloads: frameworks, force fields and adsorbates of the five benchmark configurations.

Every choice is recorded here and seeded, so that the same inputs are produced on any machine.  The generated
frameworks are cached as ``.npz`` files under ``raspa3_amd/data`` (committed), and regenerated from the seeds if
the cache is absent.

Construction (configurations 0 and 1, and the per-system frameworks of 2 and 4):
  * Si on a jittered lattice: site (i+1/2, j+1/2, k+1/2)/n in fractional coordinates, mapped to Cartesian, plus
    a uniform jitter of +-0.25 A per Cartesian axis (numpy PCG64, seed ``seed_si``).
  * Two O per Si at 1.6 A along random directions (seed ``seed_o``, batches of 256 directions); a candidate is
    rejected if it comes closer than ``d_oo`` to an O already placed or closer than 1.2 A to any other Si
    (minimum image), and of the admissible directions of a batch the one farthest from the other Si is kept.  If some Si cannot
    place its O within the attempt budget, the whole O placement restarts with the next seed; the seed actually
    used is stored with the framework.
  * Configuration 0: orthorhombic 20.0 x 19.9 x 13.4 A, 4 x 4 x 6 Si (96 Si + 192 O), d_oo = 2.4 A.
  * Configuration 1: a = b = 40, c = 30 A, gamma = 100 deg; 13 x 15 x 7 Si = 1365 Si + 2730 O = 4095 sites, exactly
    charge neutral with Si +2.05 e and O -1.025 e (4096 sites cannot be split 1:2).  d_oo = 2.0 A: 2730 O in
    47,271 A^3 (0.0578 A^-3) lies above the random-sequential-addition jamming density of 2.4 A spheres
    (about 0.053 A^-3), so 2.4 A cannot be reached by this construction.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def cell_from_lengths_angles(a, b, c, alpha=90.0, beta=90.0, gamma=90.0):
    """Cell matrix with lattice vectors as columns, as SimulationBox(a, b, c, alpha, beta, gamma) builds it
    (src/raspakit/simulationbox.cpp:40-60).  Returns (cell 3x3, is_rectangular)."""
    al, be, ga = (math.pi * x / 180.0 for x in (alpha, beta, gamma))
    rect = abs(al - 0.5 * math.pi) < 1e-5 and abs(be - 0.5 * math.pi) < 1e-5 and abs(ga - 0.5 * math.pi) < 1e-5
    t = (math.cos(al) - math.cos(ga) * math.cos(be)) / math.sin(ga)
    H = np.zeros((3, 3))
    H[:, 0] = (a, 0.0, 0.0)
    H[:, 1] = (b * math.cos(ga), b * math.sin(ga), 0.0)
    H[:, 2] = (c * math.cos(be), c * t, c * math.sqrt(1.0 - math.cos(be) ** 2 - t * t))
    return H, rect


def cell_orthorhombic(a, b, c):
    """SimulationBox(a, b, c): exactly diagonal (src/raspakit/simulationbox.cpp:15-37)."""
    return np.diag([a, b, c]).astype(float), True


@dataclass
class Framework:
    name: str
    cell: np.ndarray  # 3x3, columns = lattice vectors
    rect: bool
    pos: np.ndarray  # (n, 3)
    type: np.ndarray  # (n,) pseudo-atom index
    charge: np.ndarray  # (n,)
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.type)

    def cell9(self):
        return np.ascontiguousarray(self.cell.T.reshape(9))  # columns consecutive


def _min_image_dist(p, X, H, Hinv):
    if len(X) == 0:
        return np.full(p.shape[0], np.inf)
    d = X[None, :, :] - p[:, None, :]
    s = d @ Hinv.T
    s -= np.rint(s)
    d = s @ H.T
    return np.sqrt((d * d).sum(-1)).min(1)


def generate_silicate(name, H, rect, nsi, d_oo, seed_si, seed_o, type_si, type_o, q_si, q_o, jitter=0.25,
                      d_sio=1.6, d_osi=1.2, batch=256, max_batches=16, max_restarts=400):
    """Jittered Si lattice + two O per Si; see the module docstring."""
    Hinv = np.linalg.inv(H)
    rng = np.random.default_rng(seed_si)
    si = []
    for k in range(nsi[2]):
        for j in range(nsi[1]):
            for i in range(nsi[0]):
                f = np.array([(i + 0.5) / nsi[0], (j + 0.5) / nsi[1], (k + 0.5) / nsi[2]])
                si.append(H @ f + rng.uniform(-jitter, jitter, 3))
    si = np.array(si)
    nsi_tot = len(si)
    for restart in range(max_restarts):
        seed = seed_o + restart
        rng = np.random.default_rng(seed)
        O = np.zeros((0, 3))
        ok_all = True
        for n in range(nsi_tot):
            others = np.delete(si, n, 0)
            for _ in range(2):
                placed = False
                for _b in range(max_batches):
                    v = rng.normal(size=(batch, 3))
                    v /= np.linalg.norm(v, axis=1, keepdims=True)
                    p = si[n] + d_sio * v
                    dSi = _min_image_dist(p, others, H, Hinv)
                    good = (_min_image_dist(p, O, H, Hinv) >= d_oo) & (dSi >= d_osi)
                    idx = np.flatnonzero(good)
                    if len(idx):
                        # of the admissible directions, keep the one farthest from the other Si (the O stays
                        # inside its own Si's Voronoi cell and leaves room for the neighbours' O)
                        O = np.vstack([O, p[idx[np.argmax(dSi[idx])]]])
                        placed = True
                        break
                if not placed:
                    ok_all = False
                    break
            if not ok_all:
                break
        if ok_all:
            pos = np.vstack([si, O])
            typ = np.array([type_si] * nsi_tot + [type_o] * len(O), dtype=np.int32)
            chg = np.array([q_si] * nsi_tot + [q_o] * len(O), dtype=float)
            meta = dict(seed_si=seed_si, seed_o_used=seed, d_oo=d_oo, nsi=list(nsi), jitter=jitter)
            return Framework(name, H, rect, pos, typ, chg, meta)
    raise RuntimeError(f"{name}: O placement failed after {max_restarts} restarts")


def _cached(name, make):
    path = os.path.join(DATA_DIR, name + ".npz")
    if os.path.exists(path):
        z = np.load(path, allow_pickle=False)
        return Framework(name, z["cell"], bool(z["rect"]), z["pos"], z["type"], z["charge"],
                         json.loads(str(z["meta"])))
    fw = make()
    os.makedirs(DATA_DIR, exist_ok=True)
    np.savez(path, cell=fw.cell, rect=fw.rect, pos=fw.pos, type=fw.type, charge=fw.charge,
             meta=json.dumps(fw.meta))
    return fw


# ---------------------------------------------------------------------------------------------------------------
# Force fields.  Pseudo-atom tables as (name, eps_K, sigma_A, charge, is_framework).
# ---------------------------------------------------------------------------------------------------------------

@dataclass
class ForceFieldSpec:
    names: list
    eps_K: np.ndarray
    sigma: np.ndarray
    charge: np.ndarray
    is_framework: np.ndarray
    shifted: bool = True
    rc_framework: float = 12.0
    rc_molecule: float = 12.0
    rc_coulomb: float = 12.0
    alpha: float = 0.265
    use_charge: bool = True
    overlap: float = 1e6

    def index(self, name):
        return self.names.index(name)


def _ff(rows, **kw):
    names = [r[0] for r in rows]
    return ForceFieldSpec(names, np.array([r[1] for r in rows], float), np.array([r[2] for r in rows], float),
                          np.array([r[3] for r in rows], float), np.array([int(r[4]) for r in rows], np.int32), **kw)


def forcefield_config0():
    """LJ on O only (44 K, 3.3 A), Si without LJ; united-atom CH4 (148 K, 3.73 A); LB mixing, 12 A, shifted."""
    return _ff([("Si", 0.0, 1.0, 0.0, True), ("O", 44.0, 3.3, 0.0, True), ("CH4", 148.0, 3.73, 0.0, False)],
               use_charge=False)


def forcefield_config1():
    """Charged silicate (Si +2.05, O -1.025) with CO2 (C 27 K 2.80 A +0.70; O 79 K 3.05 A -0.35); alpha 0.265."""
    return _ff([("Si", 0.0, 1.0, 2.05, True), ("O", 44.0, 3.3, -1.025, True),
                ("C_co2", 27.0, 2.80, 0.70, False), ("O_co2", 79.0, 3.05, -0.35, False)],
               alpha=0.265, use_charge=True)


def framework_config0():
    H, rect = cell_orthorhombic(20.0, 19.9, 13.4)
    return _cached("config0_framework", lambda: generate_silicate(
        "config0_framework", H, rect, (6, 4, 4), 2.4, 1000, 1001, 0, 1, 0.0, 0.0))


def framework_config1():
    H, rect = cell_from_lengths_angles(40.0, 40.0, 30.0, 90.0, 90.0, 100.0)
    return _cached("config1_framework", lambda: generate_silicate(
        "config1_framework", H, rect, (13, 15, 7), 2.0, 1000, 1001, 0, 1, 2.05, -1.025))


def framework_small_charged():
    """A 25 A cubic charged silicate for fast Monte Carlo tests: 7 x 7 x 7 Si + 686 O = 1029 sites, neutral."""
    H, rect = cell_orthorhombic(25.0, 25.0, 25.0)
    return _cached("small_charged_framework", lambda: generate_silicate(
        "small_charged_framework", H, rect, (7, 7, 7), 2.2, 2000, 2001, 0, 1, 2.05, -1.025))


def co2_component(spec, mol_fraction=1.0):
    """Rigid CO2 of configuration 1 (C 27 K 2.80 A +0.70; O 79 K 3.05 A -0.35; C-O 1.16 A); first bead = C."""
    return dict(n_atoms=3, first_bead=1, position=[[0, 0, 1.16], [0, 0, 0], [0, 0, -1.16]],
                type=[spec.index("O_co2"), spec.index("C_co2"), spec.index("O_co2")], charge=[-0.35, 0.70, -0.35],
                mol_fraction=mol_fraction)


def n2_component(spec, mol_fraction=1.0):
    """N2 of the CO2/N2 mixture workload: two LJ sites (36 K, 3.31 A, -0.482 e) 0.55 A either side of a +0.964 e
    centre.  The parameters are the TraPPE N2 model (Potoff and Siepmann, AIChE J. 47 (2001) 1676) that the workload
    prescribes; the geometry is that of examples/basic/2_mc_co2_n2_in_two_independent_boxes/N2.json, whose force field
    uses the Martin-Calvo parameters instead (38.298 K, 3.306 A, -0.405 / +0.810 e), so the mixture results are for
    the TraPPE model, not that example's.  First bead = atom 0, an N site: the reference's
    default CBMC starting bead (src/raspakit/component.ixx:195).  Growing from the charge-only centre instead is
    exact too but samples badly: a centre trial beside a framework O scores about -87,000 K on its own (the N sites
    that keep the whole molecule away are placed only in the orientation step), so a few first-bead weights of
    order e^290 at 300 K dominate the Rosenbluth factors."""
    return dict(n_atoms=3, first_bead=0, position=[[0, 0, 0.55], [0, 0, 0], [0, 0, -0.55]],
                type=[spec.index("N_n2"), spec.index("N_com"), spec.index("N_n2")], charge=[-0.482, 0.964, -0.482],
                mol_fraction=mol_fraction)


def forcefield_co2_n2():
    """Configuration-1 silicate + CO2 + N2 (configuration 4)."""
    return _ff([("Si", 0.0, 1.0, 2.05, True), ("O", 44.0, 3.3, -1.025, True),
                ("C_co2", 27.0, 2.80, 0.70, False), ("O_co2", 79.0, 3.05, -0.35, False),
                ("N_n2", 36.0, 3.31, -0.482, False), ("N_com", 0.0, 1.0, 0.964, False)],
               alpha=0.265, use_charge=True)


def forcefield_zeolite(extra=()):
    """ForceField::makeZeoliteForceField (src/raspakit/forcefield/forcefield.cpp:1777-1841): Si 22 K 2.30 A +2.05,
    O 53 K 3.30 A -1.025, CH4 158.5 K 3.72 A, C_co2 29.933 K 2.745 A +0.6512, O_co2 85.671 K 3.017 A -0.3256;
    Ewald alpha from the automatic rule (ForceField::initializeEwaldParameters) for precision 1e-6 and 12 A.
    'extra' appends (name, eps_K, sigma_A, charge) rows."""
    rows = [("Si", 22.0, 2.30, 2.05, True), ("O", 53.0, 3.3, -1.025, True), ("CH4", 158.5, 3.72, 0.0, False),
            ("C_co2", 29.933, 2.745, 0.6512, False), ("O_co2", 85.671, 3.017, -0.3256, False)]
    rows += [(n, e, s_, q, False) for n, e, s_, q in extra]
    eps, rc = 1e-6, 12.0
    tol = math.sqrt(abs(math.log(eps * rc)))
    alpha = math.sqrt(abs(math.log(eps * rc * tol))) / rc
    return _ff(rows, alpha=alpha, use_charge=True)


def framework_itq29(spec, rep=3):
    """ITQ-29 (LTA, cubic a = 11.8671 A, 24 Si + 48 O per cell; the reference's test structure) repeated rep^3 times
    (or rep = (nx, ny, nz) times), with the force field's charges.  Returns (framework, unit cell)."""
    with open(os.path.join(DATA_DIR, "itq29.json")) as f:
        d = json.load(f)
    a = d["a"]
    frac = np.array(d["fractional"], float)
    typ = np.array([spec.index(e) for e in d["element"]], np.int32)
    n = (rep, rep, rep) if np.isscalar(rep) else tuple(rep)
    shifts = np.array([[i, j, k] for i in range(n[0]) for j in range(n[1]) for k in range(n[2])], float)
    pos = ((frac[None, :, :] + shifts[:, None, :]) * a).reshape(-1, 3)
    t = np.tile(typ, len(shifts))
    name = f"itq29x{n[0]}x{n[1]}x{n[2]}"
    return (Framework(name, np.diag([a * n[0], a * n[1], a * n[2]]), True, pos, t, spec.charge[t].astype(float)),
            np.diag([a, a, a]))


def framework_mfi(spec, rep=(2, 2, 2)):
    """MFI (silicalite-1: orthorhombic 20.022 x 19.899 x 13.383 A, 96 Si + 192 O per cell; a test structure of the
    reference, Framework::makeMFI) repeated rep = (nx, ny, nz) times, with the force field's charges.  Returns
    (framework, unit cell)."""
    with open(os.path.join(DATA_DIR, "mfi.json")) as f:
        d = json.load(f)
    abc = np.array([d["a"], d["b"], d["c"]])
    frac = np.array(d["fractional"], float)
    typ = np.array([spec.index(e) for e in d["element"]], np.int32)
    n = tuple(rep)
    shifts = np.array([[i, j, k] for i in range(n[0]) for j in range(n[1]) for k in range(n[2])], float)
    pos = ((frac[None, :, :] + shifts[:, None, :]) * abc).reshape(-1, 3)
    t = np.tile(typ, len(shifts))
    return (Framework("mfix%dx%dx%d" % n, np.diag(abc * np.array(n)), True, pos, t, spec.charge[t].astype(float)),
            np.diag(abc))


def framework_channels(spec):
    """A porous seeded silicate: a dense jittered-lattice silicate (orthorhombic 40 x 40 x 30 A, 13 x 13 x 7 Si, two O
    per Si at 1.6 A, O-O >= 2.0 A, the zeolite force field's types and charges) with two straight channels of radius
    5 A carved along c at (10, 10) and (30, 30) A.  Atoms within the radius are removed, then the remaining atoms
    nearest the axes of whichever type is in excess until twice as many O as Si are gone, so the framework stays
    neutral.  About 5 % of random points have a negative C_co2 energy (against about a third in ITQ-29)."""
    def make():
        H = np.diag([40.0, 40.0, 30.0])
        iSi, iO = spec.index("Si"), spec.index("O")
        base = generate_silicate("channels_framework", H, True, (13, 13, 7), 2.0, 2000, 2001, iSi, iO,
                                 float(spec.charge[iSi]), float(spec.charge[iO]))
        pos, typ = base.pos, base.type
        dmin = np.full(len(pos), np.inf)
        for x0, y0 in ((10.0, 10.0), (30.0, 30.0)):
            dx = pos[:, 0] - x0
            dx -= 40.0 * np.rint(dx / 40.0)
            dy = pos[:, 1] - y0
            dy -= 40.0 * np.rint(dy / 40.0)
            dmin = np.minimum(dmin, np.hypot(dx, dy))
        keep = dmin >= 5.0
        order = np.argsort(dmin, kind="stable")
        while True:
            rsi = int(((~keep) & (typ == iSi)).sum())
            ro = int(((~keep) & (typ == iO)).sum())
            if ro == 2 * rsi:
                break
            want = iO if ro < 2 * rsi else iSi
            keep[next(i for i in order if keep[i] and typ[i] == want)] = False
        meta = dict(base.meta, channels=[[10.0, 10.0], [30.0, 30.0]], channel_radius=5.0)
        return Framework("channels_framework", H, True, pos[keep], typ[keep], base.charge[keep], meta)
    return _cached("channels_framework", make)


def co2_reference(spec, mol_fraction=1.0):
    """Component::makeCO2 (src/raspakit/component.cpp:2304-2329): O (0, 0, 1.149), C (0, 0, 0), O (0, 0, -1.149);
    first bead = C."""
    return dict(n_atoms=3, first_bead=1, position=[[0, 0, 1.149], [0, 0, 0], [0, 0, -1.149]],
                type=[spec.index("O_co2"), spec.index("C_co2"), spec.index("O_co2")],
                charge=[spec.charge[spec.index("O_co2")], spec.charge[spec.index("C_co2")],
                        spec.charge[spec.index("O_co2")]], mol_fraction=mol_fraction)


def forcefield_hexane_zeolite():
    """The zeolite force field plus TraPPE united-atom CH3 (98 K, 3.75 A) and CH2 (46 K, 3.95 A)."""
    return forcefield_zeolite([("CH3", 98.0, 3.75, 0.0), ("CH2", 46.0, 3.95, 0.0)])


def hexane_component(spec, log_ideal_rosenbluth=0.0, mol_fraction=1.0):
    """United-atom n-hexane (examples/non_basic/8_mc_adsorption_c6_isomers_in_mfi/hexane.json): CH3-(CH2)4-CH3, bonds
    1.54 A, bend 114 deg with k = 62500 K/rad^2, torsion c0 + c1 (1 + cos p) + c2 (1 - cos 2p) + c3 (1 + cos 3p) with
    [0, 355.03, -68.19, 791.32] K (trans at 180 deg), LJ between beads more than three bonds apart.  Grown from
    bead 0.  log_ideal_rosenbluth: log of the ideal-gas Rosenbluth weight of the growth scheme."""
    t3, t2 = spec.index("CH3"), spec.index("CH2")
    return dict(n_atoms=6, first_bead=0, position=[[0.0, 0.0, 1.54 * i] for i in range(6)],
                type=[t3, t2, t2, t2, t2, t3], charge=[0.0] * 6, mol_fraction=mol_fraction, flexible=True,
                bond_length=1.54, bend_theta0_deg=114.0, bend_k_K=62500.0, torsion_K=[0.0, 355.03, -68.19, 791.32],
                log_ideal_rosenbluth=log_ideal_rosenbluth)


def methane_component(spec):
    return dict(n_atoms=1, first_bead=0, position=[[0, 0, 0]], type=[spec.index("CH4")], charge=[0.0],
                mol_fraction=1.0)


def mc_params(temperature=300.0, pressure_pa=1e5, seed=42, k_first=10, k_orient=10, p_translation=0.25,
              p_rotation=0.25, max_translation=0.5, max_rotation=0.5, min_steps_per_cycle=20, max_molecules=256,
              production_cycles=1000, p_reinsertion=0.0):
    return dict(temperature=temperature, pressure_pa=pressure_pa, k_first=k_first, k_orient=k_orient,
                p_translation=p_translation, p_rotation=p_rotation, max_translation=max_translation,
                max_rotation=max_rotation, min_steps_per_cycle=min_steps_per_cycle, max_molecules=max_molecules,
                seed=seed, production_cycles=production_cycles, p_reinsertion=p_reinsertion)


def grid_dims(cell, spacing):
    """n = (size_t)(w / h + 0.5) from perpendicular widths (src/raspakit/system_interpolation_grids.cpp:222-226)."""
    a, b, c = cell[:, 0], cell[:, 1], cell[:, 2]
    c1, c2, c3 = np.cross(a, b), np.cross(b, c), np.cross(a, c)
    v = abs(a @ c2)
    w = np.array([v / math.sqrt(c2 @ c2), v / math.sqrt(c3 @ c3), v / math.sqrt(c1 @ c1)])
    return [int(x / spacing + 0.5) for x in w], w

"""raspa3_amd: RASPA3's framework–adsorbate energy path on MI355X (gfx950).

Python host layer over the C ABI (include/raspa3_amd/rg.h), used by the tests and the benchmark.  All compute runs
in libraspa3_amd.so on the GPU; this module only moves host arrays in and out.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib as L
from ._lib import RGError, lib

# Internal units of the reference (src/foundationkit/units.ixx:25-60, src/raspakit/units.cpp:15-100): length 1 A,
# time 1 ps, mass 1 u; energies in 10 J/mol.
_E_UNIT = 1.6605402e-27 * 1e-10 * 1e-10 / (1e-12 * 1e-12)
ENERGY_TO_KELVIN = _E_UNIT * 6.0221419947e23 / 8.314464919
KB = 1.380650324e-23 / _E_UNIT           # Boltzmann constant in internal energy units per K
PRESSURE_TO_PA = 1.6605402e-27 / (1e-10 * 1e-12 * 1e-12)


def beta(temperature):
    """1 / (kB T) in inverse internal energy units."""
    return 1.0 / (KB * temperature)


__all__ = ["Context", "ForceField", "Framework", "Grid", "box", "grid_dims", "RGError", "lib"]


def box(a, b, c, alpha=90.0, beta=90.0, gamma=90.0, lengths_only=False):
    """SimulationBox (src/raspakit/simulationbox.cpp:15-60): returns dict(cell, inv, rect, volume, widths)."""
    cell = np.zeros(9)
    inv = np.zeros(9)
    rect = C.c_int()
    vol = C.c_double()
    w = np.zeros(3)
    L.check(lib().rg_box(a, b, c, alpha, beta, gamma, int(lengths_only), L.d(cell), L.d(inv), C.byref(rect),
                         C.byref(vol), L.d(w)), None, "rg_box")
    return dict(cell=cell.reshape(3, 3).T.copy(), inv=inv.reshape(3, 3).T.copy(), rect=bool(rect.value),
                volume=vol.value, widths=w)


def grid_dims(cell, spacing):
    n = np.zeros(3, np.int32)
    c9 = L.f64(np.asarray(cell).T.reshape(9))
    L.check(lib().rg_grid_dims(L.d(c9), spacing, L.i(n)), None, "rg_grid_dims")
    return [int(x) for x in n]


class Context:
    """rg_ctx: one HIP device and stream.  Closing it first closes every object still open on it (newest first),
    so that nothing is released after its context."""

    def __init__(self, device=0):
        h = C.c_void_p()
        st = lib().rg_ctx_create(device, C.byref(h))
        if st != L.RG_OK:
            raise RGError(f"rg_ctx_create failed with status {st} (no usable HIP device?)")
        self.h = h
        self._children = []

    def _own(self, obj):
        self._children.append(weakref.ref(obj))

    def close(self):
        if self.h:
            for r in reversed(self._children):
                obj = r()
                if obj is not None:
                    obj.close()
            self._children = []
            lib().rg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        L.check(lib().rg_sync(self.h), self.h, "rg_sync")

    @property
    def stream(self):
        return lib().rg_stream(self.h)

    def error(self):
        return lib().rg_last_error(self.h).decode()


class _Handle:
    """A library object owned by a Context; 'close' releases it once (also on garbage collection)."""

    _destroy = None
    h = None

    def _register(self, ctx, h):
        self.h = h
        if ctx is not None:
            ctx._own(self)

    def close(self):
        if self.h:
            getattr(lib(), self._destroy)(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ForceField(_Handle):
    """rg_forcefield built from a workloads.ForceFieldSpec (or any object with the same fields)."""

    _destroy = "rg_forcefield_destroy"

    def __init__(self, ctx, spec):
        self.ctx = ctx
        self.spec = spec
        h = C.c_void_p()
        eps, sig, isf = L.f64(spec.eps_K), L.f64(spec.sigma), L.i32(spec.is_framework)
        L.check(lib().rg_forcefield_create(ctx.h if ctx else None, len(eps), L.d(eps), L.d(sig), L.i(isf),
                                           int(spec.shifted), spec.rc_framework, spec.rc_molecule, spec.rc_coulomb,
                                           spec.alpha, int(spec.use_charge), spec.overlap, C.byref(h)),
                ctx.h if ctx else None, "rg_forcefield_create")
        self._register(ctx, h)
        self.n = len(eps)

    def pairs(self):
        n = self.n
        eps, sig, sh = np.zeros(n * n), np.zeros(n * n), np.zeros(n * n)
        lj = np.zeros(n * n, np.int32)
        L.check(lib().rg_forcefield_pairs(self.h, L.d(eps), L.d(sig), L.d(sh), L.i(lj)), None, "pairs")
        return eps.reshape(n, n), sig.reshape(n, n), sh.reshape(n, n), lj.reshape(n, n)



class Framework(_Handle):
    """rg_framework from a workloads.Framework."""

    _destroy = "rg_framework_destroy"

    def __init__(self, ctx, fw):
        self.ctx = ctx
        self.fw = fw
        h = C.c_void_p()
        c9 = L.f64(fw.cell.T.reshape(9))
        pos, typ, q = L.f64(fw.pos), L.i32(fw.type), L.f64(fw.charge)
        L.check(lib().rg_framework_create(ctx.h, L.d(c9), int(fw.rect), len(typ), L.d(pos), L.i(typ), L.d(q),
                                          C.byref(h)), ctx.h, "rg_framework_create")
        self._register(ctx, h)


class Grid(_Handle):
    """rg_grid: a device-resident interpolation grid, layout_left (values, nx, ny, nz); scheme 8 = tricubic,
    27 = triquintic (ForceField::InterpolationScheme)."""

    _destroy = "rg_grid_destroy"

    def __init__(self, ctx, cell, rect, n, scheme=8):
        self.ctx = ctx
        self.cell = np.asarray(cell, float)
        self.rect = bool(rect)
        self.n = tuple(int(x) for x in n)
        self.values = int(scheme)
        h = C.c_void_p()
        c9 = L.f64(self.cell.T.reshape(9))
        L.check(lib().rg_grid_create_scheme(ctx.h, L.d(c9), int(rect), *self.n, self.values, C.byref(h)), ctx.h,
                "rg_grid_create_scheme")
        self._register(ctx, h)

    @property
    def count(self):
        return self.values * self.n[0] * self.n[1] * self.n[2]

    def download(self, offset=0, count=None):
        count = self.count - offset if count is None else count
        out = np.empty(count)
        L.check(lib().rg_grid_download(self.ctx.h, self.h, offset, count, L.d(out)), self.ctx.h, "download")
        return out

    def upload(self, data, offset=0):
        data = L.f64(data).reshape(-1)
        L.check(lib().rg_grid_upload(self.ctx.h, self.h, offset, data.size, L.d(data)), self.ctx.h, "upload")

    def interpolate(self, pos, gradient=False):
        pos = L.f64(pos).reshape(-1, 3)
        e = np.empty(len(pos))
        g = np.empty((len(pos), 3)) if gradient else None
        L.check(lib().rg_grid_interpolate(self.ctx.h, self.h, len(pos), L.d(pos), L.d(e), L.d(g)), self.ctx.h,
                "rg_grid_interpolate")
        return (e, g) if gradient else e

    def interpolate_hessian(self, pos):
        """InterpolationEnergyGrid::interpolateHessian: energies (m,), gradients (m, 3), Hessians (m, 3, 3)."""
        pos = L.f64(pos).reshape(-1, 3)
        e = np.empty(len(pos))
        g = np.empty((len(pos), 3))
        h = np.empty((len(pos), 3, 3))
        L.check(lib().rg_grid_interpolate_hessian(self.ctx.h, self.h, len(pos), L.d(pos), L.d(e), L.d(g), L.d(h)),
                self.ctx.h, "rg_grid_interpolate_hessian")
        return e, g, h



def ewald_parameters(precision, rc, widths):
    """ForceField::initializeEwaldParameters (automatic): returns (alpha, kmax[3])."""
    a = C.c_double()
    k = np.zeros(3, np.int32)
    L.check(lib().rg_ewald_parameters(precision, rc, L.d(L.f64(widths)), C.byref(a), L.i(k)), None,
            "rg_ewald_parameters")
    return a.value, [int(x) for x in k]


class System(_Handle):
    """rg_system: box, force field, framework (direct) and/or grids, adsorbates, Ewald state."""

    _destroy = "rg_system_destroy"

    def __init__(self, ctx, ff, cell, rect, framework=None, grids=None, coulomb_grid=None, kmax=(8, 8, 8),
                 int_cut2=0, capacity=4096):
        self.ctx = ctx
        self._keep = (ff, framework, grids, coulomb_grid)
        grids = grids or {}
        gtypes = L.i32(list(grids.keys()))
        garr = (C.c_void_p * max(1, len(grids)))(*[g.h.value for g in grids.values()])
        h = C.c_void_p()
        c9 = L.f64(np.asarray(cell).T.reshape(9))
        L.check(lib().rg_system_create(ctx.h, ff.h, L.d(c9), int(rect), framework.h if framework else None,
                                       len(grids), L.i(gtypes), garr, coulomb_grid.h if coulomb_grid else None,
                                       L.i(L.i32(kmax)), int(int_cut2), int(capacity), C.byref(h)),
                ctx.h, "rg_system_create")
        self._register(ctx, h)

    def info(self):
        nk, n = C.c_int(), C.c_int()
        si, q = C.c_double(), C.c_double()
        L.check(lib().rg_system_info(self.h, C.byref(nk), C.byref(si), C.byref(q), C.byref(n)), None, "info")
        return dict(nk=nk.value, single_ion_sum=si.value, net_charge=q.value, n_adsorbate_atoms=n.value)

    def set_adsorbates(self, pos, type, charge, molecule):
        pos, t, q, m = L.f64(pos).reshape(-1, 3), L.i32(type), L.f64(charge), L.i32(molecule)
        L.check(lib().rg_system_set_adsorbates(self.ctx.h, self.h, len(t), L.d(pos), L.i(t), L.d(q), L.i(m)),
                self.ctx.h, "rg_system_set_adsorbates")

    def trial_energies(self, pos, type, charge, molecule, skip_molecule=-1):
        pos, t, q, m = L.f64(pos).reshape(-1, 3), L.i32(type), L.f64(charge), L.i32(molecule)
        e = np.zeros((len(t), 4))
        ov = np.zeros(len(t), np.int32)
        L.check(lib().rg_trial_energies(self.ctx.h, self.h, len(t), L.d(pos), L.i(t), L.d(q), L.i(m), skip_molecule,
                                        L.d(e), L.i(ov)), self.ctx.h, "rg_trial_energies")
        return e, ov.astype(bool)

    def ewald_delta(self, new_pos, new_q, new_mol, old_pos, old_q, old_mol):
        def arr(p, q, m):
            return L.f64(p).reshape(-1, 3), L.f64(q), L.i32(m)
        npos, nq, nm = arr(new_pos, new_q, new_mol)
        opos, oq, om = arr(old_pos, old_q, old_mol)
        out = np.zeros(3)
        L.check(lib().rg_ewald_delta(self.ctx.h, self.h, len(nq), L.d(npos), L.d(nq), L.i(nm), len(oq), L.d(opos),
                                     L.d(oq), L.i(om), L.d(out)), self.ctx.h, "rg_ewald_delta")
        return out

    def ewald_accept(self):
        L.check(lib().rg_ewald_accept(self.ctx.h, self.h), self.ctx.h, "rg_ewald_accept")

    def ewald_energy(self):
        out = np.zeros(3)
        L.check(lib().rg_ewald_energy(self.ctx.h, self.h, L.d(out)), self.ctx.h, "rg_ewald_energy")
        return out



def rosenbluth_select(ctx, energies, overlap, beta, uniforms):
    """energies, overlap: (n_sets, k).  Returns (selected[n_sets], log_weight[n_sets])."""
    U = L.f64(energies)
    ns, k = U.shape
    ov = L.i32(overlap)
    u = L.f64(uniforms)
    sel = np.zeros(ns, np.int32)
    lw = np.zeros(ns)
    L.check(lib().rg_rosenbluth_select(ctx.h, ns, k, L.d(U), L.i(ov), beta, L.d(u), L.i(sel), L.d(lw)), ctx.h,
            "rg_rosenbluth_select")
    return sel, lw


def build_grids(ctx, ff, fw, specs, sync=True, split_radius=None):
    """Builds several equal-dimension grids in one fused pass.  specs: list of (Grid, type, probe, cutoff).
    split_radius: None for FP64 pair terms throughout, else the radius beyond which the Ewald real-space pair terms are
    computed in FP32 (rg_grid_build_many_precision)."""
    n = len(specs)
    types = L.i32([s[1] for s in specs])
    probes = L.i32([s[2] for s in specs])
    cuts = L.f64([s[3] for s in specs])
    arr = (C.c_void_p * n)(*[s[0].h.value for s in specs])
    if split_radius is not None:
        L.check(lib().rg_grid_build_many_precision(ctx.h, ff.h, fw.h, n, L.i(types), L.i(probes), L.d(cuts), arr,
                                                   L.RG_GRID_SPLIT_FP32, float(split_radius), int(not sync)),
                ctx.h, "rg_grid_build_many_precision")
        return
    fn = lib().rg_grid_build_many if sync else lib().rg_grid_build_many_async
    L.check(fn(ctx.h, ff.h, fw.h, n, L.i(types), L.i(probes), L.d(cuts), arr), ctx.h, "rg_grid_build_many")


def grid_key(ff, fw, grid_type, probe, cutoff, split_radius=None):
    """rg_grid_key: 64-bit key of the inputs a grid is built from (the check of a grid cache file)."""
    k = C.c_ulonglong()
    prec = L.RG_GRID_FP64 if split_radius is None else L.RG_GRID_SPLIT_FP32
    L.check(lib().rg_grid_key(ff.h, fw.h, int(grid_type), int(probe), float(cutoff), prec,
                              0.0 if split_radius is None else float(split_radius), C.byref(k)), None, "rg_grid_key")
    return k.value


def grid_save(ctx, grid, key, path):
    L.check(lib().rg_grid_save(ctx.h, grid.h, C.c_ulonglong(key), str(path).encode()), ctx.h, "rg_grid_save")


def grid_load(ctx, grid, key, path):
    L.check(lib().rg_grid_load(ctx.h, grid.h, C.c_ulonglong(key), str(path).encode()), ctx.h, "rg_grid_load")


def grid_write_cube(ctx, grid, path):
    L.check(lib().rg_grid_write_cube(ctx.h, grid.h, str(path).encode()), ctx.h, "rg_grid_write_cube")


def grid_check(ctx, ff, fw, lj_grid, probe, coulomb_grid=None, probe_charge=0.0, temperature=300.0,
               n_points=100000, frac=None, seed=1):
    """rg_grid_check: the reference's test of a built grid on random points (Boltzmann-weighted averages of the
    interpolated and analytic energy, gradient and, for triquintic grids, Hessian; relative errors).  Returns dict(n_points, lennard_jones,
    n_coulomb_points, coulomb)."""
    out = L.GridCheck()
    f = None if frac is None else L.f64(frac).reshape(-1, 3)
    n = n_points if f is None else len(f)
    L.check(lib().rg_grid_check(ctx.h, ff.h, fw.h, lj_grid.h, int(probe),
                                None if coulomb_grid is None else coulomb_grid.h, float(probe_charge),
                                float(temperature), int(n), L.d(f), C.c_ulonglong(seed), C.byref(out)), ctx.h,
            "rg_grid_check")

    def stats(x):
        return dict(weight=x.weight, energy_table=x.energy_table, energy_full=x.energy_full,
                    energy_relative_error=x.energy_relative_error, gradient_table=np.array(list(x.gradient_table)),
                    gradient_full=np.array(list(x.gradient_full)),
                    gradient_relative_error=np.array(list(x.gradient_relative_error)),
                    hessian_table=np.array(list(x.hessian_table)), hessian_full=np.array(list(x.hessian_full)),
                    hessian_relative_error=np.array(list(x.hessian_relative_error)))

    return dict(n_points=out.n_points, lennard_jones=stats(out.lennard_jones), n_coulomb_points=out.n_coulomb_points,
                coulomb=stats(out.coulomb))


def block_error(block_averages, average):
    """rg_block_error: 95 % confidence interval of block averages (the reference's blockErrorEstimate)."""
    b = L.f64(block_averages)
    ci = C.c_double()
    L.check(lib().rg_block_error(len(b), L.d(b), float(average), C.byref(ci)), None, "rg_block_error")
    return ci.value


def block_average(sums, weights):
    """rg_block_average: (mean, ci) of a directly averaged property from per-block weighted sums and weights."""
    s, w = L.f64(sums), L.f64(weights)
    mean, ci = C.c_double(), C.c_double()
    L.check(lib().rg_block_average(len(s), L.d(s), L.d(w), C.byref(mean), C.byref(ci)), None, "rg_block_average")
    return mean.value, ci.value


def enthalpy_of_adsorption(weight, sum_u, sum_n, sum_un, sum_nn, temperature):
    """rg_enthalpy_of_adsorption: per-component enthalpy of adsorption (internal units; the isosteric heat is minus
    it) and its 95 % block interval, from per-block sums shaped [b], [b], [b, c], [b, c], [b, c, c]."""
    w = L.f64(weight)
    nb = len(w)
    n = L.f64(sum_n).reshape(nb, -1)
    nc = n.shape[1]
    mean, ci = np.zeros(nc), np.zeros(nc)
    L.check(lib().rg_enthalpy_of_adsorption(nb, nc, L.d(w), L.d(L.f64(sum_u)), L.d(n), L.d(L.f64(sum_un)),
                                            L.d(L.f64(sum_nn)), float(temperature), L.d(mean), L.d(ci)), None,
            "rg_enthalpy_of_adsorption")
    return mean, ci


def widom_statistics(weight, sum_w, sum_density, temperature):
    """rg_widom_statistics: Rosenbluth weight, excess / ideal-gas / total chemical potential and fugacity with their
    95 % block intervals (internal units)."""
    w = L.f64(weight)
    out = L.WidomStats()
    L.check(lib().rg_widom_statistics(len(w), L.d(w), L.d(L.f64(sum_w)), L.d(L.f64(sum_density)), float(temperature),
                                      C.byref(out)), None, "rg_widom_statistics")
    return {n: getattr(out, n) for n, _ in L.WidomStats._fields_}


def mc_statistics(state, n_components, temperature):
    """Loading per component (mean and 95 % block interval) and enthalpy of adsorption per component of one system
    from its rg_mc_state_t block sums, as the reference reports them."""
    w = state["block_count"]
    nc = n_components
    loading = [block_average(state["block_n"][:, c], w) for c in range(nc)]
    dh, dh_ci = enthalpy_of_adsorption(w, state["block_energy"], state["block_n"][:, :nc],
                                       state["block_energy_n"][:, :nc], state["block_nn"][:, :nc, :nc], temperature)
    return dict(loading=np.array([m for m, _ in loading]), loading_ci=np.array([c for _, c in loading]),
                enthalpy=dh, enthalpy_ci=dh_ci)


def component_desc(d):
    """rg_component_desc from a dict(n_atoms, first_bead, position (n,3), type, charge, mol_fraction[, flexible,
    bond_length, bend_theta0_deg, bend_k_K, torsion_K, log_ideal_rosenbluth])."""
    c = L.ComponentDesc()
    c.n_atoms = d["n_atoms"]
    c.first_bead = d["first_bead"]
    pos = np.asarray(d["position"], float).reshape(-1)
    for k in range(len(pos)):
        c.position[k] = pos[k]
    for a in range(d["n_atoms"]):
        c.type[a] = int(d["type"][a])
        c.charge[a] = float(d["charge"][a])
    c.mol_fraction = d.get("mol_fraction", 1.0)
    if d.get("flexible", False):
        c.flexible = 1
        c.bond_length = d["bond_length"]
        c.bend_theta0_deg = d["bend_theta0_deg"]
        c.bend_k_K = d["bend_k_K"]
        for k in range(4):
            c.torsion_K[k] = d["torsion_K"][k]
        c.log_ideal_rosenbluth = d.get("log_ideal_rosenbluth", 0.0)
    return c


def mc_params(p):
    q = L.McParams()
    for name, _ in L.McParams._fields_:
        setattr(q, name, p[name])
    return q


def widom(ctx, system, comp, params, first, count):
    """rg_widom: returns (log_weight, energy, ewald) per insertion."""
    lw, u, ew = np.zeros(count), np.zeros(count), np.zeros(count)
    c = component_desc(comp)
    p = mc_params(params)
    L.check(lib().rg_widom(ctx.h, system.h, C.byref(c), C.byref(p), int(first), int(count), L.d(lw), L.d(u), L.d(ew)),
            ctx.h, "rg_widom")
    return lw, u, ew


def widom_device(ctx, system, comp, params, first, count, log_weight, energy, ewald):
    """rg_widom_device: the same insertions with device-resident outputs (addresses of count doubles each, e.g.
    torch tensors' data_ptr()); returns when the kernel has finished."""
    c = component_desc(comp)
    p = mc_params(params)
    L.check(lib().rg_widom_device(ctx.h, system.h, C.byref(c), C.byref(p), int(first), int(count),
                                  C.c_void_p(int(log_weight)), C.c_void_p(int(energy)), C.c_void_p(int(ewald))),
            ctx.h, "rg_widom_device")


def autotune_summary():
    """Launch configurations chosen by the library's online autotuner, one line per tuned problem."""
    n = lib().rg_autotune_summary(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().rg_autotune_summary(buf, n + 1)
    return buf.value.decode()


def molecule_weights(ctx, system, comp, temperature, frac, n_orient=16, seed=1):
    """log of the orientation-averaged Boltzmann factor of a rigid molecule with its first bead at each point
    (fractional coordinates of the system's box), framework energy from the grids."""
    frac = L.f64(frac).reshape(-1, 3)
    out = np.empty(len(frac))
    desc = component_desc(comp)
    L.check(lib().rg_molecule_weights(ctx.h, system.h, C.byref(desc), temperature, len(frac), L.d(frac), n_orient,
                                      C.c_ulonglong(seed), L.d(out)), ctx.h, "rg_molecule_weights")
    return out


class Mc(_Handle):
    """rg_mc: GCMC over several systems (one wavefront each)."""

    _destroy = "rg_mc_destroy"

    def __init__(self, ctx, systems, comps, params):
        self.ctx = ctx
        self.systems = systems
        arr = (C.c_void_p * len(systems))(*[s.h.value for s in systems])
        descs = (L.ComponentDesc * len(comps))(*[component_desc(c) for c in comps])
        self.slots = max(int(c["n_atoms"]) for c in comps)  # atom slots per molecule in the adsorbate arrays
        p = mc_params(params)
        h = C.c_void_p()
        L.check(lib().rg_mc_create(ctx.h, len(systems), arr, len(comps), descs, C.byref(p), C.byref(h)), ctx.h,
                "rg_mc_create")
        self._register(ctx, h)

    def set_proposal(self, system, n3, prob, component=None):
        """First-bead proposal of a system: for every component, or for one (component index)."""
        n = L.i32(n3)
        p = None if prob is None else L.f64(prob).reshape(-1)
        if component is None:
            L.check(lib().rg_mc_set_proposal(self.ctx.h, self.h, system, L.i(n), L.d(p)), self.ctx.h, "set_proposal")
        else:
            L.check(lib().rg_mc_set_component_proposal(self.ctx.h, self.h, system, int(component), L.i(n), L.d(p)),
                    self.ctx.h, "set_component_proposal")

    def set_launch_shape(self, shape):
        """-1: autotuned; 0, 1, 2: one wavefront per system, 1, 2 or 4 systems per workgroup; 3: four per system."""
        L.check(lib().rg_mc_set_launch_shape(self.ctx.h, self.h, int(shape)), self.ctx.h, "rg_mc_set_launch_shape")

    def set_balance(self, enable=True, overhead_molecules=256.0):
        """Balanced launches: every system runs at least one cycle per launch and further whole cycles while their
        cost (steps x (molecules + overhead_molecules)) stays within the largest single-cycle cost of the launch."""
        L.check(lib().rg_mc_set_balance(self.ctx.h, self.h, int(bool(enable)), float(overhead_molecules)),
                self.ctx.h, "rg_mc_set_balance")

    def run(self, cycles, sample=False, graph=True):
        L.check(lib().rg_mc_run(self.ctx.h, self.h, int(cycles), int(sample), int(graph)), self.ctx.h, "rg_mc_run")

    def state(self, system):
        s = L.McState()
        L.check(lib().rg_mc_state(self.ctx.h, self.h, system, C.byref(s)), self.ctx.h, "rg_mc_state")
        return dict(n_molecules=s.n_molecules, n_per_component=list(s.n_per_component), energy=s.energy,
                    steps=s.steps, cycles=s.cycles_sampled, trials=np.array(list(s.trials)),
                    accepted=np.array(list(s.accepted)),
                    block_count=np.array(list(s.block_count)), block_energy=np.array(list(s.block_energy)),
                    block_n=np.array([list(r) for r in s.block_n]),
                    block_energy_n=np.array([list(r) for r in s.block_energy_n]),
                    block_nn=np.array([[list(q) for q in r] for r in s.block_nn]))

    def adsorbates(self, system, capacity=1 << 16):
        pos = np.zeros((capacity, 3))
        t = np.zeros(capacity, np.int32)
        c = np.zeros(capacity, np.int32)
        n = C.c_int()
        L.check(lib().rg_mc_adsorbates(self.ctx.h, self.h, system, capacity, L.d(pos), L.i(t), L.i(c), C.byref(n)),
                self.ctx.h, "rg_mc_adsorbates")
        return pos[:n.value], t[:n.value], c[:n.value // self.slots]

    def total_energy(self):
        out = np.zeros(len(self.systems))
        L.check(lib().rg_mc_total_energy(self.ctx.h, self.h, L.d(out)), self.ctx.h, "rg_mc_total_energy")
        return out



class Mlp(_Handle):
    """rg_mlp: the learned first-bead proposal network (BF16 MFMA training and inference on the device)."""

    _destroy = "rg_mlp_destroy"

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        h = C.c_void_p()
        L.check(lib().rg_mlp_create(ctx.h, C.c_ulonglong(seed), C.byref(h)), ctx.h, "rg_mlp_create")
        self._register(ctx, h)
        self.n_params = int(lib().rg_mlp_num_params())

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        L.check(lib().rg_mlp_get_params(self.ctx.h, self.h, L.f32p(out)), self.ctx.h, "rg_mlp_get_params")
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32).reshape(-1)
        assert w.size == self.n_params
        L.check(lib().rg_mlp_set_params(self.ctx.h, self.h, L.f32p(w)), self.ctx.h, "rg_mlp_set_params")

    def forward(self, frac):
        frac = L.f64(frac).reshape(-1, 3)
        out = np.zeros(len(frac))
        L.check(lib().rg_mlp_forward(self.ctx.h, self.h, len(frac), L.d(frac), L.d(out)), self.ctx.h,
                "rg_mlp_forward")
        return out

    def gradients(self, frac, energy, beta):
        frac = L.f64(frac).reshape(-1, 3)
        energy = L.f64(energy).reshape(-1)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double()
        L.check(lib().rg_mlp_gradients(self.ctx.h, self.h, len(frac), L.d(frac), L.d(energy), beta, L.f32p(g),
                                       C.byref(loss)), self.ctx.h, "rg_mlp_gradients")
        return g, loss.value

    def train(self, frac, energy, beta, batch=4096, epochs=1, lr=1e-3, seed=1):
        frac = L.f64(frac).reshape(-1, 3)
        energy = L.f64(energy).reshape(-1)
        losses = np.zeros(max(epochs, 1))
        L.check(lib().rg_mlp_train(self.ctx.h, self.h, len(frac), L.d(frac), L.d(energy), beta, batch, epochs, lr,
                                   C.c_ulonglong(seed), L.d(losses)), self.ctx.h, "rg_mlp_train")
        return losses[:epochs]

    def proposal(self, n3, floor=0.05):
        n = L.i32(n3)
        out = np.zeros(int(np.prod(n3)))
        L.check(lib().rg_mlp_proposal(self.ctx.h, self.h, L.i(n), floor, L.d(out)), self.ctx.h, "rg_mlp_proposal")
        return out


"""ctypes binding of the CPU oracle (oracle/gale_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module.  Nothing under ``gale_amd/`` imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgale_oracle.so")

YIELD_NONE, YIELD_DRUCKER_PRAGER, YIELD_VON_MISES = 0, 1, 2
OK, E_NAN, E_MAXIT, E_ZERO_FORCE, E_INVALID, E_ESCAPED, E_INNER_MAXIT, E_PICARD = 0, 1, 2, 3, 6, 7, 8, 9


class Material(C.Structure):
    _fields_ = [
        ("eta0", C.c_double), ("density", C.c_double), ("yield_type", C.c_int32), ("pad_", C.c_int32),
        ("cohesion", C.c_double), ("friction", C.c_double), ("min_yield_stress", C.c_double),
        ("min_viscosity", C.c_double), ("max_strain_rate", C.c_double),
    ]


class ProblemDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("res", C.c_int32 * 3), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("bc_mask", C.c_void_p), ("bc_value", C.c_void_p), ("gravity", C.c_double), ("ghat", C.c_double * 3),
        ("eta_clamp_min", C.c_double), ("eta_clamp_max", C.c_double), ("n_materials", C.c_int32),
        ("pad_", C.c_int32), ("materials", C.c_void_p), ("n_particles", C.c_int64), ("coord", C.c_void_p),
        ("material", C.c_void_p), ("eta0_particle", C.c_void_p),
    ]


class UzawaParams(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double), ("max_iterations", C.c_int32), ("min_iterations", C.c_int32),
        ("use_absolute_tolerance", C.c_int32), ("use_preconditioner", C.c_int32), ("inner", C.c_int32),
        ("inner_max_iterations", C.c_int32), ("inner_rtol_factor", C.c_double),
        ("inner_rtol_override", C.c_double), ("nthreads", C.c_int32), ("pad_", C.c_int32),
    ]


class PicardParams(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double), ("max_iterations", C.c_int32), ("min_iterations", C.c_int32),
        ("kill_non_convergent", C.c_int32), ("pad_", C.c_int32),
    ]


class Report(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("uzawa_iterations", C.c_int32), ("inner_iterations", C.c_int64),
        ("rel_residual", C.c_double), ("abs_residual", C.c_double), ("div_u", C.c_double),
        ("momentum_residual", C.c_double), ("solve_seconds", C.c_double), ("assemble_seconds", C.c_double),
        ("picard_iterations", C.c_int32), ("pad_", C.c_int32), ("picard_residual", C.c_double),
    ]


class Rng(C.Structure):
    _fields_ = [("state", C.c_long)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        sig = {
            "orc_rng_seed": (None, [vp, C.c_long]),
            "orc_rng_next": (d, [vp]),
            "orc_rng_range": (d, [vp, d, d]),
            "orc_vertex_coord": (d, [i32, d, d, i32]),
            "orc_isotropic_D": (None, [i32, d, vp]),
            "orc_isotropic_stress": (None, [i32, d, vp, vp]),
            "orc_search_element": (i32, [i32, vp, vp, vp, vp, vp]),
            "orc_layout_random": (i64, [i32, vp, vp, vp, i32, C.c_long, vp]),
            "orc_sobol_sequence": (i32, [C.c_uint, C.c_uint, vp, i64, vp, vp]),
            "orc_initial_damage": (i32, [i64, vp, vp, C.c_long, vp, vp]),
            "orc_layout_spacefiller": (i64, [i32, vp, vp, vp, i64, d, vp, vp, vp, vp]),
            "orc_problem_new": (vp, [vp]),
            "orc_problem_free": (None, [vp]),
            "orc_num_nodes": (i64, [vp]),
            "orc_num_elements": (i64, [vp]),
            "orc_num_particles": (i64, [vp]),
            "orc_get_particles": (None, [vp, vp, vp, vp, vp]),
            "orc_set_coords": (i32, [vp, vp]),
            "orc_particle_viscosity": (i32, [vp, vp, vp, i32, vp, vp]),
            "orc_element_K": (i32, [vp, i64, vp]),
            "orc_element_G": (i32, [vp, i64, vp]),
            "orc_element_f": (i32, [vp, i64, vp]),
            "orc_assemble": (i32, [vp, vp, vp, i32, i32]),
            "orc_system_sizes": (None, [vp, vp]),
            "orc_system_export": (None, [vp] * 11),
            "orc_apply_K_full": (None, [vp, vp, vp]),
            "orc_apply_G_full": (None, [vp, vp, vp]),
            "orc_apply_Gt_full": (None, [vp, vp, vp]),
            "orc_diag_K_full": (None, [vp, vp]),
            "orc_uzawa_Q": (None, [vp, vp]),
            "orc_uzawa": (i32, [vp, vp, vp, vp, vp, vp]),
            "orc_picard": (i32, [vp, vp, vp, vp, vp, vp]),
            "orc_picard_from": (i32, [vp, vp, vp, vp, vp, vp, vp, vp]),
            "orc_compute_dt": (d, [vp, vp, d]),
            "orc_interpolate_velocity": (i32, [vp, vp, vp, vp]),
            "orc_advect_rk2": (i32, [vp, vp, d, vp]),
            "orc_dvc_cell": (i32, [i32, vp, i32, vp, vp, vp]),
            "orc_dvc_cell_owners": (i32, [i32, vp, i32, vp, vp, vp, vp]),
            "orc_set_weights": (i32, [vp, i32, vp]),
            "orc_set_strain_weakening": (i32, [vp, i32, vp]),
            "orc_set_plastic_strain": (i32, [vp, vp]),
            "orc_plastic_strain": (i32, [vp, vp, vp]),
            "orc_weakening_step": (i32, [vp, d]),
            "orc_weakening_step_limited": (i32, [vp, d, vp, d]),
            "orc_particle_fields": (i32, [vp, vp, vp, vp, vp, vp]),
            "orc_population_control": (i32, [vp, vp, vp]),
            "orc_get_cell_tables": (None, [vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class RngStream:
    """Park-Miller stream with the reference's seed mask (Random.cxx:45-74)."""

    def __init__(self, seed: int):
        self._g = Rng()
        lib().orc_rng_seed(C.byref(self._g), seed)

    def next(self) -> float:
        return lib().orc_rng_next(C.byref(self._g))

    def range(self, mn: float, mx: float) -> float:
        return lib().orc_rng_range(C.byref(self._g), mn, mx)


def isotropic_D(dim, eta):
    nst = 3 if dim == 2 else 6
    D = np.zeros((nst, nst))
    lib().orc_isotropic_D(dim, eta, _p(D))
    return D


def isotropic_stress(dim, eta, strain_rate):
    sr = np.ascontiguousarray(strain_rate, dtype=np.float64)
    st = np.zeros_like(sr)
    lib().orc_isotropic_stress(dim, eta, _p(sr), _p(st))
    return st


def vertex_coord(n, lo, hi, idx):
    return lib().orc_vertex_coord(n, lo, hi, idx)


def search_element(dim, res, lo, hi, x):
    r = np.asarray(list(res) + [1] * (3 - len(res)), dtype=np.int32)
    l = np.asarray(list(lo) + [0.0] * (3 - len(lo)), dtype=np.float64)
    h = np.asarray(list(hi) + [1.0] * (3 - len(hi)), dtype=np.float64)
    xx = np.asarray(list(x) + [0.0] * (3 - len(x)), dtype=np.float64)
    out = np.zeros(1, dtype=np.int32)
    ok = lib().orc_search_element(dim, _p(r), _p(l), _p(h), _p(xx), _p(out))
    return int(out[0]) if ok else -1


def layout_random(dim, res, lo, hi, ppc, seed):
    r = np.asarray(list(res) + [1] * (3 - len(res)), dtype=np.int32)
    l = np.asarray(list(lo) + [0.0] * (3 - len(lo)), dtype=np.float64)
    h = np.asarray(list(hi) + [1.0] * (3 - len(hi)), dtype=np.float64)
    E = int(np.prod(r[:dim]))
    coord = np.zeros((E * ppc, 3), dtype=np.float64)
    lib().orc_layout_random(dim, _p(r), _p(l), _p(h), ppc, seed, _p(coord))
    return coord


def sobol_sequence(degree, coefficient, m, count):
    """count numbers of one Sobol generator (SobolGenerator.cxx:198-294):
    (by GetNextNumber, by GetNumberByIndex(k))."""
    mm = np.zeros(32, dtype=np.uint32)
    mm[:len(m)] = m
    nxt = np.zeros(count, dtype=np.float64)
    idx = np.zeros(count, dtype=np.float64)
    if lib().orc_sobol_sequence(degree, coefficient, _p(mm), count, _p(nxt), _p(idx)) != 0:
        raise ValueError("sobol_sequence")
    return nxt, idx


def layout_spacefiller(dim, res, lo, hi, generators, total=0, avg_per_cell=0.0):
    """SpaceFillerParticleLayout.cxx:211-240: particles at the Sobol points of
    `generators` (one (degree, coefficient, m) per axis) scaled to [lo, hi]."""
    r = np.asarray(list(res) + [1] * (3 - len(res)), dtype=np.int32)
    l = np.asarray(list(lo) + [0.0] * (3 - len(lo)), dtype=np.float64)
    h = np.asarray(list(hi) + [1.0] * (3 - len(hi)), dtype=np.float64)
    deg = np.zeros(3, dtype=np.uint32)
    coef = np.zeros(3, dtype=np.uint32)
    m = np.zeros((3, 9), dtype=np.uint32)
    for d, (dg, cf, mv) in enumerate(generators[:dim]):
        deg[d], coef[d] = dg, cf
        m[d, :len(mv)] = mv
    E = int(np.prod(r[:dim]))
    P = int(total) if total > 0 else int(np.uint32(E * avg_per_cell))
    coord = np.zeros((P, 3), dtype=np.float64)
    n = lib().orc_layout_spacefiller(dim, _p(r), _p(l), _p(h), int(total), float(avg_per_cell), _p(deg), _p(coef),
                                     _p(m), _p(coord))
    assert n == P
    return coord


def initial_damage(coord, fraction, final_softening_strain, seed, factor=1.0, wavenumber=-1.0, sin_i=-1.0,
                   cos_i=-1.0, sin_k=-1.0, cos_k=-1.0, inside=None):
    """StrainWeakening.cxx:333-430 initial damage, drawing from the C library's rand()."""
    c = np.ascontiguousarray(coord, dtype=np.float64).reshape(-1, 3)
    prm = np.array([fraction, factor, final_softening_strain, wavenumber, sin_i, cos_i, sin_k, cos_k], np.float64)
    ins = None if inside is None else np.ascontiguousarray(inside, dtype=np.uint8)
    w = np.zeros(len(c))
    lib().orc_initial_damage(len(c), _p(c), _p(prm), int(seed), None if ins is None else _p(ins), _p(w))
    return w


def dvc_cell(dim, res, xi):
    """Discrete Voronoi centroids and weights of one cell's particles
    (DVCWeights.cxx:693-786 growth, :202-258 centroids); xi: (n, dim) local
    coordinates in cell-table order, res: sub-grid per axis."""
    xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, dim)
    n = xi.shape[0]
    r = np.array(list(res) + [1] * (3 - len(res)), dtype=np.int32)
    cent = np.zeros((n, dim))
    w = np.zeros(n)
    st = lib().orc_dvc_cell(dim, r.ctypes.data, n, xi.ctypes.data, cent.ctypes.data, w.ctypes.data)
    if st != 0:
        raise RuntimeError(f"orc_dvc_cell status {st}")
    return cent, w


def dvc_cell_owners(dim, res, xi):
    """dvc_cell plus the owning particle of every sub-cell (k = i + r0 (j + r1 l))."""
    xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, dim)
    n = xi.shape[0]
    r = np.array(list(res) + [1] * (3 - len(res)), dtype=np.int32)
    cent = np.zeros((n, dim))
    w = np.zeros(n)
    own = np.zeros(int(np.prod(r)), np.int32)
    st = lib().orc_dvc_cell_owners(dim, r.ctypes.data, n, xi.ctypes.data, cent.ctypes.data, w.ctypes.data,
                                   own.ctypes.data)
    if st != 0:
        raise RuntimeError(f"orc_dvc_cell status {st}")
    return cent, w, own


def uzawa_params(tolerance=1e-5, max_iterations=1000, min_iterations=1, use_absolute_tolerance=0,
                 use_preconditioner=1, inner=0, inner_max_iterations=0, inner_rtol_factor=0.1,
                 inner_rtol_override=0.0, nthreads=1):
    return UzawaParams(tolerance, max_iterations, min_iterations, use_absolute_tolerance, use_preconditioner,
                       inner, inner_max_iterations, inner_rtol_factor, inner_rtol_override, nthreads, 0)


class Problem:
    """A Q1-P0 PIC Stokes problem held by the oracle."""

    def __init__(self, dim, res, lo, hi, coord, material, materials, bc_mask=None, bc_value=None,
                 gravity=1.0, ghat=(0.0, -1.0, 0.0), eta_clamp=(0.0, 0.0), eta0_particle=None):
        self.dim = dim
        self._keep = []
        d = ProblemDesc()
        d.dim = dim
        for i in range(3):
            d.res[i] = res[i] if i < dim else 1
            d.lo[i] = lo[i] if i < dim else 0.0
            d.hi[i] = hi[i] if i < dim else 1.0
            d.ghat[i] = ghat[i]
        self.res = tuple(d.res[:dim])
        self.N = int(np.prod([r + 1 for r in self.res]))
        self.E = int(np.prod(self.res))
        if bc_mask is not None:
            bc_mask = np.ascontiguousarray(bc_mask, dtype=np.uint8).reshape(-1)
            assert bc_mask.size == self.N * dim
            self._keep.append(bc_mask)
            d.bc_mask = _p(bc_mask)
        if bc_value is not None:
            bc_value = _f64(bc_value).reshape(-1)
            self._keep.append(bc_value)
            d.bc_value = _p(bc_value)
        d.gravity = gravity
        d.eta_clamp_min, d.eta_clamp_max = eta_clamp
        mats = (Material * len(materials))()
        for i, m in enumerate(materials):
            mats[i] = Material(m.get("eta0", 1.0), m.get("density", 0.0), m.get("yield_type", 0), 0,
                               m.get("cohesion", 0.0), m.get("friction", 0.0), m.get("min_yield_stress", 0.0),
                               m.get("min_viscosity", 0.0), m.get("max_strain_rate", 0.0))
        self._keep.append(mats)
        d.n_materials = len(materials)
        d.materials = C.addressof(mats)
        coord = _f64(coord).reshape(-1, 3)
        material = np.ascontiguousarray(material, dtype=np.int32)
        self._keep += [coord, material]
        d.n_particles = coord.shape[0]
        d.coord = _p(coord)
        d.material = _p(material)
        if eta0_particle is not None:
            e0 = _f64(eta0_particle)
            self._keep.append(e0)
            d.eta0_particle = _p(e0)
        self.P = coord.shape[0]
        self._h = lib().orc_problem_new(C.byref(d))
        if not self._h:
            raise ValueError("oracle rejected the problem (particle outside the box or bad material index)")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_problem_free(h)
            self._h = None

    def population_control(self, upperT=25.0, lowerT=0.6, maxSplits=3, maxDeletions=2,
                           splitInInterfaceCells=False, deleteInInterfaceCells=False, resolution=(10, 10, 10),
                           Inflow=False, particlesPerCell=25, Threshold=0.8, CentPosRatio=0.01, rand_seed=1,
                           rand_skip=0):
        """One PCDVC pass (the reference's XML names); returns (added,
        deleted) and sets self.rand_draws to the rand() calls it made."""
        class _P(C.Structure):
            _fields_ = [("upper_t", C.c_double), ("lower_t", C.c_double), ("max_splits", C.c_int32),
                        ("max_deletions", C.c_int32), ("split_in_interface_cells", C.c_int32),
                        ("delete_in_interface_cells", C.c_int32), ("res", C.c_int32 * 3), ("inflow", C.c_int32),
                        ("particles_per_cell", C.c_int32), ("pad_", C.c_int32), ("threshold", C.c_double),
                        ("cent_pos_ratio", C.c_double), ("rand_seed", C.c_int64), ("rand_skip", C.c_int64)]
        r = list(resolution) + [10] * (3 - len(resolution))
        p = _P(upperT, lowerT, maxSplits, maxDeletions, int(splitInInterfaceCells), int(deleteInInterfaceCells),
               (C.c_int32 * 3)(*r), int(Inflow), particlesPerCell, 0, Threshold, CentPosRatio, rand_seed, rand_skip)
        counts = np.zeros(3, np.int64)
        st = lib().orc_population_control(self._h, C.byref(p), _p(counts))
        if st != 0:
            raise RuntimeError(f"orc_population_control status {st}")
        self.P = int(lib().orc_num_particles(self._h))
        self.rand_draws = int(counts[2])
        return int(counts[0]), int(counts[1])

    def particle_fields(self, u_full):
        """Strain rate, its invariant, viscosity and stress per material point
        (caller order) for the full velocity u_full; the viscosity is the one
        of the last constitutive pass (uzawa / assemble)."""
        nst = 3 if self.dim == 2 else 6
        u = _f64(u_full)
        sr = np.zeros((self.P, nst))
        inv = np.zeros(self.P)
        eta = np.zeros(self.P)
        st = np.zeros((self.P, nst))
        code = lib().orc_particle_fields(self._h, _p(u), _p(sr), _p(inv), _p(eta), _p(st))
        if code != 0:
            raise RuntimeError(f"orc_particle_fields status {code}")
        return dict(strain_rate=sr, strain_rate_inv=inv, viscosity=eta, stress=st)

    def set_weights(self, kind, res=(10, 10, 10)):
        """0: ConstantWeights; 1: DVCWeights with sub-grid res."""
        r = np.array(list(res) + [10] * (3 - len(res)), dtype=np.int32)
        st = lib().orc_set_weights(self._h, kind, r.ctypes.data)
        if st != 0:
            raise RuntimeError(f"orc_set_weights status {st}")

    def set_strain_weakening(self, material, healing_rate=0.0, initial_softening_strain=0.0,
                             final_softening_strain=float("inf"), cohesion_after=0.0, friction_after=0.0):
        """StrainWeakening on `material`'s yield rheology (None-like: pass
        healing_rate=None to detach)."""
        if healing_rate is None:
            st = lib().orc_set_strain_weakening(self._h, material, None)
        else:
            sw = np.array([healing_rate, initial_softening_strain, final_softening_strain, cohesion_after,
                           friction_after], dtype=np.float64)
            st = lib().orc_set_strain_weakening(self._h, material, sw.ctypes.data)
        if st != 0:
            raise RuntimeError(f"orc_set_strain_weakening status {st}")

    def set_plastic_strain(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.shape == (self.P,)
        if lib().orc_set_plastic_strain(self._h, w.ctypes.data) != 0:
            raise RuntimeError("orc_set_plastic_strain")

    def plastic_strain(self):
        """(w, increment) in caller order."""
        w = np.zeros(self.P)
        inc = np.zeros(self.P)
        if lib().orc_plastic_strain(self._h, w.ctypes.data, inc.ctypes.data) != 0:
            raise RuntimeError("orc_plastic_strain")
        return w, inc

    def weakening_step(self, dt, inside=None, limit=0.0):
        """inside: optional (P,) flags of the strain-limited region (positions after the step)."""
        ins = None if inside is None else np.ascontiguousarray(inside, dtype=np.uint8)
        if lib().orc_weakening_step_limited(self._h, float(dt), _p(ins), float(limit)) != 0:
            raise RuntimeError("orc_weakening_step")

    def particles(self):
        owner = np.zeros(self.P, dtype=np.int32)
        xi = np.zeros((self.P, self.dim))
        w = np.zeros(self.P)
        coord = np.zeros((self.P, 3))
        lib().orc_get_particles(self._h, _p(owner), _p(xi), _p(w), _p(coord))
        return owner, xi, w, coord

    def cell_tables(self):
        """Each cell's particles (caller indices) in the cell table's order."""
        off = np.zeros(int(lib().orc_num_elements(self._h)) + 1, np.int64)
        lst = np.zeros(max(self.P, 1), np.int64)
        lib().orc_get_cell_tables(self._h, _p(off), _p(lst))
        return [lst[off[e]:off[e + 1]].copy() for e in range(len(off) - 1)]

    def set_coords(self, coord):
        c = _f64(coord).reshape(-1, 3)
        return lib().orc_set_coords(self._h, _p(c))

    def viscosity(self, u_prev=None, p_prev=None, yield_on=0):
        eta = np.zeros(self.P)
        yl = np.zeros(self.P, dtype=np.uint8)
        u_prev, p_prev = _f64(u_prev), _f64(p_prev)
        st = lib().orc_particle_viscosity(self._h, _p(u_prev), _p(p_prev), yield_on, _p(eta), _p(yl))
        return st, eta, yl

    def element_K(self, e):
        nd = (4 if self.dim == 2 else 8) * self.dim
        Ke = np.zeros((nd, nd))
        lib().orc_element_K(self._h, e, _p(Ke))
        return Ke

    def element_G(self, e):
        nd = (4 if self.dim == 2 else 8) * self.dim
        Ge = np.zeros(nd)
        lib().orc_element_G(self._h, e, _p(Ge))
        return Ge

    def element_f(self, e):
        nd = (4 if self.dim == 2 else 8) * self.dim
        fe = np.zeros(nd)
        lib().orc_element_f(self._h, e, _p(fe))
        return fe

    def assemble(self, u_prev=None, p_prev=None, yield_on=0, nthreads=1):
        u_prev, p_prev = _f64(u_prev), _f64(p_prev)
        return lib().orc_assemble(self._h, _p(u_prev), _p(p_prev), yield_on, nthreads)

    def system(self):
        """Reduced system (K_ff, G_fp, f, h, Q, eq) as scipy CSR matrices and arrays."""
        import scipy.sparse as sp
        s = np.zeros(4, dtype=np.int64)
        lib().orc_system_sizes(self._h, _p(s))
        nu, npr, nk, ng = (int(v) for v in s)
        Kp, Kc, Kv = np.zeros(nu + 1, np.int64), np.zeros(nk, np.int32), np.zeros(nk)
        Gp, Gc, Gv = np.zeros(nu + 1, np.int64), np.zeros(ng, np.int32), np.zeros(ng)
        f, h, Q = np.zeros(nu), np.zeros(npr), np.zeros(npr)
        eq = np.zeros(self.N * self.dim, np.int32)
        lib().orc_system_export(self._h, _p(Kp), _p(Kc), _p(Kv), _p(Gp), _p(Gc), _p(Gv), _p(f), _p(h), _p(Q),
                                _p(eq))
        K = sp.csr_matrix((Kv, Kc, Kp), shape=(nu, nu))
        G = sp.csr_matrix((Gv, Gc, Gp), shape=(nu, npr))
        return K, G, f, h, Q, eq

    def apply_K(self, x):
        x = _f64(x).reshape(-1)
        y = np.zeros_like(x)
        lib().orc_apply_K_full(self._h, _p(x), _p(y))
        return y

    def apply_G(self, p):
        p = _f64(p)
        y = np.zeros(self.N * self.dim)
        lib().orc_apply_G_full(self._h, _p(p), _p(y))
        return y

    def apply_Gt(self, x):
        x = _f64(x).reshape(-1)
        q = np.zeros(self.E)
        lib().orc_apply_Gt_full(self._h, _p(x), _p(q))
        return q

    def diag_K(self):
        d = np.zeros(self.N * self.dim)
        lib().orc_diag_K_full(self._h, _p(d))
        return d

    def uzawa_Q(self):
        q = np.zeros(self.E)
        lib().orc_uzawa_Q(self._h, _p(q))
        return q

    def uzawa(self, params=None, p0=None, **kw):
        params = params or uzawa_params(**kw)
        u = np.zeros(self.N * self.dim)
        p = np.zeros(self.E)
        rep = Report()
        p0 = _f64(p0)
        st = lib().orc_uzawa(self._h, C.byref(params), _p(p0), _p(u), _p(p), C.byref(rep))
        return st, u, p, rep

    def picard(self, params=None, tolerance=0.01, max_iterations=500, min_iterations=1, kill=0, u_prev=None,
               p_prev=None, **kw):
        """Picard loop; u_prev (full nodal) / p_prev: the previous solution
        (a previous time step's, or a restart's): the first solve then yields
        from it and starts from p_prev."""
        params = params or uzawa_params(**kw)
        pp = PicardParams(tolerance, max_iterations, min_iterations, kill, 0)
        u = np.zeros(self.N * self.dim)
        p = np.zeros(self.E)
        rep = Report()
        u_prev, p_prev = _f64(u_prev), _f64(p_prev)
        st = lib().orc_picard_from(self._h, C.byref(params), C.byref(pp), _p(u_prev), _p(p_prev), _p(u), _p(p),
                                   C.byref(rep))
        return st, u, p, rep

    def compute_dt(self, u, dt_factor=1.0):
        u = _f64(u).reshape(-1)
        return lib().orc_compute_dt(self._h, _p(u), dt_factor)

    def interpolate(self, u, x):
        u = _f64(u).reshape(-1)
        xx = np.asarray(list(x) + [0.0] * (3 - len(x)), dtype=np.float64)
        v = np.zeros(3)
        ok = lib().orc_interpolate_velocity(self._h, _p(u), _p(xx), _p(v))
        return (v[: self.dim] if ok else None)

    def advect_rk2(self, u, dt):
        u = _f64(u).reshape(-1)
        out = np.zeros((self.P, 3))
        st = lib().orc_advect_rk2(self._h, _p(u), dt, _p(out))
        return st, out

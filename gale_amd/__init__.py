"""gale_amd -- MI355X Q1-P0 PIC Stokes hot path behind the reference's
Stokes_SLE / Stokes_SLE_UzawaSolver surface.

The compute path is ``libgale_stokes.so`` (hand-written HIP for gfx950, C ABI
declared in ``include/gale_stokes.h`` and ``include/gale_generate.h``).  This
module binds that library with ctypes and mirrors the reference's component
interface (XML parameter names, Solve / GetRhs / FormResidual, Picard) so that
callers and tests read like the reference's own.  There is no CPU fallback:
if the library or a GPU is missing, calls raise ``GaleError``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgale_stokes.so")

OK, E_NAN, E_MAXIT, E_ZERO_FORCE, E_STAGNATION, E_HIP, E_INVALID, E_ESCAPED, E_INNER_MAXIT, E_PICARD, E_NO_DEVICE = range(11)
YIELD_NONE, YIELD_DRUCKER_PRAGER, YIELD_VON_MISES = 0, 1, 2
PC_NONE, PC_DIAGONAL, PC_LEARNED = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
INNER_JACOBI, INNER_MULTIGRID = 0, 1
APPLY_PARTICLES_FP64, APPLY_PARTICLES_FP32, APPLY_MOMENTS_FP64, APPLY_MOMENTS_FP32 = 0, 1, 2, 3
APPLY_G, APPLY_GT = 4, 5
SHAPE_EVERYWHERE, SHAPE_SPHERE, SHAPE_BOX, SHAPE_BELOW_SINES = 0, 1, 2, 3
WALLS = {"left": 0, "right": 1, "bottom": 2, "top": 3, "back": 4, "front": 5}

# Integration weights of the material points (include/gale_weights.h).
WEIGHTS_CONSTANT, WEIGHTS_DVC = 0, 1

# Exported entry points (must match the declarations of include/*.h).
EXPORTS = (
    "gs_strerror", "gs_version", "gs_device_count", "gs_problem_create", "gs_problem_destroy", "gs_problem_info",
    "gs_problem_particles", "gs_solver_create", "gs_solver_set_params", "gs_solver_destroy", "gs_solver_setup",
    "gs_stokes_solve", "gs_solver_set_pressure", "gs_solver_set_solution", "gs_picard_solve", "gs_get_rhs", "gs_form_residual",
    "gs_compute_dt", "gs_advect_rk2", "gs_relocate", "gs_problem_set_coords", "gs_apply_K", "gs_apply_K_fp32",
    "gs_apply_G", "gs_apply_Gt", "gs_apply_K_noalloc", "gs_solver_sync", "gs_time_apply_K", "gs_time_apply", "gs_solver_setup_arrays", "gs_apply_K_moments",
    "gs_rng_stream", "gs_layout_random", "gs_layout_gauss", "gs_sobol_sequence", "gs_layout_spacefiller", "gs_crand_stream", "gs_initial_damage", "gs_wall_bc",
    "gs_h5_dataset_info", "gs_h5_read_f64", "gs_h5_read_i32", "gs_h5_attribute_i64", "gs_h5_read_field",
    "gs_h5_read_swarm", "gs_h5_write_field", "gs_h5_write_swarm", "gs_h5_write_time",
    "gs_h5_write_swarm_vars", "gs_h5_read_swarm_var", "gs_shape_layout", "gs_gaussian_field",
    "gs_lpc_create", "gs_lpc_destroy", "gs_lpc_num_params", "gs_lpc_get_params", "gs_lpc_set_params",
    "gs_lpc_forward", "gs_lpc_backward", "gs_lpc_loss", "gs_lpc_dataset_create", "gs_lpc_dataset_destroy",
    "gs_lpc_dataset_size", "gs_lpc_dataset_get", "gs_solver_record_pairs", "gs_lpc_train", "gs_lpc_eval",
    "gs_solver_set_learned", "gs_lpc_selftest",
    "gs_batch_create", "gs_batch_destroy", "gs_batch_set_tile", "gs_batch_autotune", "gs_batch_solve",
    "gs_batch_solution",
    "gs_problem_set_weights", "gs_problem_set_strain_weakening", "gs_problem_set_plastic_strain",
    "gs_problem_set_strain_limit",
    "gs_problem_plastic_strain", "gs_dvc_cells", "gs_dvc_max_particles", "gs_solver_particle_fields",
    "gs_problem_population_control",
)


class GaleError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = _strerror(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else f"{msg} (status {status})")


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
        ("use_absolute_tolerance", C.c_int32), ("preconditioner", C.c_int32), ("precision", C.c_int32),
        ("inner_max_iterations", C.c_int32), ("inner_rtol_factor", C.c_double),
        ("inner_rtol_override", C.c_double), ("monitor", C.c_int32), ("check_every", C.c_int32),
        ("inner_preconditioner", C.c_int32), ("mg_smooth", C.c_int32), ("flexible_window", C.c_int32),
        ("stagnation_window", C.c_int32), ("learned_restarts", C.c_int32), ("learned_q_every", C.c_int32),
        ("inner_relax_max", C.c_double), ("inner_relax_gain", C.c_double), ("mg_smooth_coarse", C.c_int32),
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
        ("momentum_residual", C.c_double), ("solve_seconds", C.c_double), ("setup_seconds", C.c_double),
        ("picard_iterations", C.c_int32), ("learned_fallback", C.c_int32), ("picard_residual", C.c_double),
        ("k_applies", C.c_int64), ("verify_rounds", C.c_int32), ("pad2_", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad_"}


class Shape(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("material", C.c_int32), ("axis", C.c_int32), ("n_modes", C.c_int32),
        ("p", C.c_double * 6), ("modes", C.c_double * 32),
    ]


_lib = None


def lib():
    """Load libgale_stokes.so (built in-tree by ``gale_amd.build_lib``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GaleError(E_NO_DEVICE, f"{LIB_PATH} is not built (run python -m gale_amd.build_lib)")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        sig = {
            "gs_strerror": (C.c_char_p, [i32]),
            "gs_version": (i32, []),
            "gs_device_count": (i32, []),
            "gs_problem_create": (i32, [vp, vp]),
            "gs_problem_destroy": (None, [vp]),
            "gs_problem_info": (i32, [vp, vp]),
            "gs_problem_particles": (i32, [vp, vp, vp, vp, vp]),
            "gs_solver_create": (i32, [vp, vp, vp]),
            "gs_solver_set_params": (i32, [vp, vp]),
            "gs_solver_destroy": (None, [vp]),
            "gs_solver_setup": (i32, [vp, vp, vp, i32]),
            "gs_stokes_solve": (i32, [vp, vp, vp, vp]),
            "gs_solver_set_pressure": (i32, [vp, vp]),
            "gs_solver_set_solution": (i32, [vp, vp, vp]),
            "gs_picard_solve": (i32, [vp, vp, vp, vp, vp]),
            "gs_get_rhs": (i32, [vp, vp]),
            "gs_form_residual": (i32, [vp, vp, vp]),
            "gs_compute_dt": (i32, [vp, d, vp]),
            "gs_advect_rk2": (i32, [vp, d]),
            "gs_relocate": (i32, [vp, vp]),
            "gs_problem_set_coords": (i32, [vp, vp]),
            "gs_problem_set_weights": (i32, [vp, i32, vp]),
            "gs_problem_set_strain_weakening": (i32, [vp, i32, vp]),
            "gs_problem_set_plastic_strain": (i32, [vp, vp]),
            "gs_problem_set_strain_limit": (i32, [vp, i32, vp, d]),
            "gs_problem_plastic_strain": (i32, [vp, vp, vp]),
            "gs_dvc_cells": (i32, [i32, vp, i64, vp, vp, vp, vp]),
            "gs_dvc_max_particles": (i64, [i32, vp]),
            "gs_solver_particle_fields": (i32, [vp, vp, vp, vp, vp, vp]),
            "gs_problem_population_control": (i32, [vp, vp, vp]),
            "gs_apply_K": (i32, [vp, vp, vp]),
            "gs_apply_K_noalloc": (i32, [vp, vp, vp]),
            "gs_apply_K_fp32": (i32, [vp, vp, vp]),
            "gs_apply_K_moments": (i32, [vp, vp, vp]),
            "gs_apply_G": (i32, [vp, vp, vp]),
            "gs_apply_Gt": (i32, [vp, vp, vp]),
            "gs_solver_sync": (i32, [vp]),
            "gs_time_apply_K": (i32, [vp, i32, i32, vp]),
            "gs_time_apply": (i32, [vp, i32, i32, i32, vp]),
            "gs_solver_setup_arrays": (i32, [vp, vp, vp, vp, vp, vp]),
            "gs_rng_stream": (d, [i64, i64, vp]),
            "gs_layout_random": (i64, [i32, vp, vp, vp, i32, i64, vp]),
            "gs_layout_gauss": (i64, [i32, vp, vp, vp, vp]),
            "gs_sobol_sequence": (i32, [i32, i64, i64, vp]),
            "gs_crand_stream": (i32, [i64, i64, vp]),
            "gs_initial_damage": (i32, [i32, i64, vp, vp, vp, vp]),
            "gs_layout_spacefiller": (i64, [i32, vp, vp, vp, i64, d, i32, vp]),
            "gs_h5_dataset_info": (i32, [C.c_char_p, C.c_char_p, vp, vp, vp]),
            "gs_h5_read_f64": (i64, [C.c_char_p, C.c_char_p, vp, i64]),
            "gs_h5_read_i32": (i64, [C.c_char_p, C.c_char_p, vp, i64]),
            "gs_h5_attribute_i64": (i32, [C.c_char_p, C.c_char_p, vp, i32]),
            "gs_h5_read_field": (i64, [C.c_char_p, i32, i32, i64, vp]),
            "gs_h5_read_swarm": (i64, [C.c_char_p, C.c_char_p, i32, i64, vp, vp]),
            "gs_h5_write_field": (i32, [C.c_char_p, i32, i32, i64, vp, vp, vp]),
            "gs_h5_write_swarm": (i32, [C.c_char_p, i32, i64, vp, vp]),
            "gs_h5_write_time": (i32, [C.c_char_p, d, d, i32]),
            "gs_h5_write_swarm_vars": (i32, [C.c_char_p, i32, i64, vp, vp, i32, vp, vp]),
            "gs_h5_read_swarm_var": (i64, [C.c_char_p, C.c_char_p, C.c_char_p, i64, vp]),
            "gs_wall_bc": (i32, [i32, vp, i32, i32, d, vp, vp]),
            "gs_shape_layout": (i32, [i32, i64, vp, i32, vp, vp]),
            "gs_gaussian_field": (i32, [i32, vp, vp, d, i64, d, d, vp]),
            "gs_lpc_create": (i32, [C.c_uint64, vp]),
            "gs_lpc_destroy": (None, [vp]),
            "gs_lpc_num_params": (i64, []),
            "gs_lpc_get_params": (i32, [vp, vp]),
            "gs_lpc_set_params": (i32, [vp, vp]),
            "gs_lpc_forward": (i32, [vp, i32, i32, i32, i32, vp, vp]),
            "gs_lpc_backward": (i32, [vp, i32, i32, i32, i32, vp, vp, vp]),
            "gs_lpc_loss": (i32, [vp, i32, i32, i32, i32, vp, vp, vp, vp, vp, vp]),
            "gs_lpc_dataset_create": (i32, [i32, i32, i32, i32, vp]),
            "gs_lpc_dataset_destroy": (None, [vp]),
            "gs_lpc_dataset_size": (i32, [vp]),
            "gs_lpc_dataset_get": (i32, [vp, i32, vp, vp, vp, vp]),
            "gs_solver_record_pairs": (i32, [vp, vp, i32]),
            "gs_lpc_train": (i32, [vp, vp, vp, vp]),
            "gs_lpc_eval": (i32, [vp, vp, i32, i32, vp]),
            "gs_solver_set_learned": (i32, [vp, vp]),
            "gs_lpc_selftest": (i32, [vp]),
            "gs_batch_create": (i32, [i32, vp, vp]),
            "gs_batch_destroy": (None, [vp]),
            "gs_batch_set_tile": (i32, [vp, i32, i32]),
            "gs_batch_autotune": (i32, [vp, vp, vp, vp, vp]),
            "gs_batch_solve": (i32, [vp, vp, vp, vp]),
            "gs_batch_solution": (i32, [vp, i32, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _strerror(status: int) -> str:
    try:
        return lib().gs_strerror(status).decode()
    except Exception:  # library missing
        return "gale_stokes library unavailable"


def _check(status: int, what: str = "", allow=()):
    if status != OK and status not in allow:
        raise GaleError(status, what)
    return status


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _dims(dim, vals, fill):
    v = list(vals)[:dim] + [fill] * (3 - min(dim, len(vals)))
    return v[:3]


def device_count() -> int:
    return int(lib().gs_device_count())


# ---------------------------------------------------------------------------
# generation helpers (gale_generate.h)
# ---------------------------------------------------------------------------
def rng_stream(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count)
    lib().gs_rng_stream(seed, count, _ptr(out))
    return out


def layout_random(dim, res, lo, hi, ppc, seed=13) -> np.ndarray:
    r = np.asarray(_dims(dim, res, 1), np.int32)
    l, h = np.asarray(_dims(dim, lo, 0.0), np.float64), np.asarray(_dims(dim, hi, 1.0), np.float64)
    E = int(np.prod(r[:dim]))
    coord = np.zeros((E * ppc, 3))
    n = lib().gs_layout_random(dim, _ptr(r), _ptr(l), _ptr(h), ppc, seed, _ptr(coord))
    if n < 0:
        raise GaleError(E_INVALID, "layout_random")
    return coord


def layout_gauss(dim, res, lo, hi) -> np.ndarray:
    r = np.asarray(_dims(dim, res, 1), np.int32)
    l, h = np.asarray(_dims(dim, lo, 0.0), np.float64), np.asarray(_dims(dim, hi, 1.0), np.float64)
    E = int(np.prod(r[:dim]))
    coord = np.zeros((E * 2 ** dim, 3))
    lib().gs_layout_gauss(dim, _ptr(r), _ptr(l), _ptr(h), _ptr(coord))
    return coord


def sobol_sequence(generator: int, count: int, first: int = 0) -> np.ndarray:
    """Numbers first .. first+count-1 of tabulated Sobol generator `generator`
    (SobolGenerator_GetNumberByIndex, StgDomain/Utils/src/SobolGenerator.cxx:271-294)."""
    out = np.zeros(count)
    if lib().gs_sobol_sequence(generator, first, count, _ptr(out)) != 0:
        raise GaleError(E_INVALID, "sobol_sequence")
    return out


def layout_spacefiller(dim, res, lo, hi, total=0, avg_per_cell=0.0, first_generator=0) -> np.ndarray:
    """SpaceFillerParticleLayout (StgDomain/Swarm/src/SpaceFillerParticleLayout.cxx:211-240):
    quasi-random particles over the whole box, totalInitialParticles = total or
    averageInitialParticlesPerCell = avg_per_cell (exactly one of the two)."""
    r = np.asarray(_dims(dim, res, 1), np.int32)
    l, h = np.asarray(_dims(dim, lo, 0.0), np.float64), np.asarray(_dims(dim, hi, 1.0), np.float64)
    n = lib().gs_layout_spacefiller(dim, _ptr(r), _ptr(l), _ptr(h), int(total), float(avg_per_cell),
                                    first_generator, None)
    if n < 0:
        raise GaleError(E_INVALID, "layout_spacefiller")
    coord = np.zeros((n, 3))
    lib().gs_layout_spacefiller(dim, _ptr(r), _ptr(l), _ptr(h), int(total), float(avg_per_cell), first_generator,
                                _ptr(coord))
    return coord


# ---------------------------------------------------------------------------
# HDF5 checkpoints (gale_checkpoint.h)
# ---------------------------------------------------------------------------
H5_I32, H5_I64, H5_F32, H5_F64 = 1, 2, 3, 4
_H5_ERRORS = {-1: "cannot read file", -2: "not an HDF5 file or corrupt structure",
              -3: "outside the supported HDF5 subset", -4: "not found", -5: "shape mismatch",
              -6: "invalid argument"}


def _h5(st, what):
    if st < 0:
        raise GaleError(E_INVALID, f"{what}: {_H5_ERRORS.get(int(st), st)}")
    return st


def h5_dataset_info(path, name):
    """(shape, type) of root-group dataset `name` (type one of H5_*)."""
    rank = np.zeros(1, np.int32)
    dims = np.zeros(4, np.int64)
    typ = np.zeros(1, np.int32)
    _h5(lib().gs_h5_dataset_info(os.fsencode(path), name.encode(), _ptr(rank), _ptr(dims), _ptr(typ)), name)
    return tuple(int(v) for v in dims[:rank[0]]), int(typ[0])


def h5_read(path, name) -> np.ndarray:
    """Root-group dataset `name` as float64 (floating types) or int32 (integer types)."""
    shape, typ = h5_dataset_info(path, name)
    n = int(np.prod(shape)) if shape else 1
    if typ in (H5_I32, H5_I64):
        out = np.zeros(n, np.int32)
        _h5(lib().gs_h5_read_i32(os.fsencode(path), name.encode(), _ptr(out), n), name)
    else:
        out = np.zeros(n, np.float64)
        _h5(lib().gs_h5_read_f64(os.fsencode(path), name.encode(), _ptr(out), n), name)
    return out.reshape(shape)


def h5_attribute(path, name, capacity=16):
    """Integer attribute `name` of the root group, or None when absent."""
    out = np.zeros(capacity, np.int64)
    n = lib().gs_h5_attribute_i64(os.fsencode(path), name.encode(), _ptr(out), capacity)
    if n == -4:
        return None
    _h5(n, name)
    return out[:min(n, capacity)].copy()


def read_field(path, dim, dofs, n_nodes) -> np.ndarray:
    """FeVariable_ReadFromFile: (n_nodes, dofs) nodal values of a field checkpoint."""
    out = np.zeros((n_nodes, dofs))
    _h5(lib().gs_h5_read_field(os.fsencode(path), dim, dofs, n_nodes, _ptr(out)), "read_field")
    return out


def read_swarm(path, swarm, dim):
    """FileParticleLayout on a SwarmDump file: (coord (P, 3), material index (P,))."""
    P = _h5(lib().gs_h5_read_swarm(os.fsencode(path), swarm.encode(), dim, 0, None, None), "read_swarm")
    coord = np.zeros((P, 3))
    mat = np.zeros(P, np.int32)
    _h5(lib().gs_h5_read_swarm(os.fsencode(path), swarm.encode(), dim, P, _ptr(coord), _ptr(mat)), "read_swarm")
    return coord, mat


def write_field(path, dim, values, coords=None, mesh_res=None):
    """FeVariable_SaveToFile, checkpoint version 2: values (n_nodes, dofs)."""
    v = np.ascontiguousarray(values, np.float64)
    v = v.reshape(v.shape[0], -1)
    c = None if coords is None else np.ascontiguousarray(coords, np.float64).reshape(-1, 3)
    r = None if mesh_res is None else np.ascontiguousarray(list(mesh_res)[:dim], np.int32)
    _h5(lib().gs_h5_write_field(os.fsencode(path), dim, v.shape[1], v.shape[0], _ptr(v),
                                None if c is None else _ptr(c), None if r is None else _ptr(r)), "write_field")


def write_swarm(path, dim, coord, material, variables=None):
    """SwarmDump of a material swarm: coordinates (P, 3), material indices (P,)
    and optional further scalar variables {name: (P,) values}."""
    c = np.ascontiguousarray(coord, np.float64).reshape(-1, 3)
    m = np.ascontiguousarray(material, np.int32).reshape(-1)
    if len(m) != len(c):
        raise GaleError(E_INVALID, "write_swarm sizes")
    if not variables:
        _h5(lib().gs_h5_write_swarm(os.fsencode(path), dim, len(c), _ptr(c), _ptr(m)), "write_swarm")
        return
    arrs = [np.ascontiguousarray(v, np.float64).reshape(-1) for v in variables.values()]
    if any(len(a) != len(c) for a in arrs):
        raise GaleError(E_INVALID, "write_swarm variable sizes")
    names = (C.c_char_p * len(arrs))(*[k.encode() for k in variables])
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    _h5(lib().gs_h5_write_swarm_vars(os.fsencode(path), dim, len(c), _ptr(c), _ptr(m), len(arrs), names, ptrs),
        "write_swarm")


def read_swarm_variable(path, swarm, name):
    """One scalar swarm variable (e.g. "PostFailureWeakening") of a SwarmDump file."""
    n = _h5(lib().gs_h5_read_swarm_var(os.fsencode(path), swarm.encode(), name.encode(), 0, None), name)
    out = np.zeros(n)
    _h5(lib().gs_h5_read_swarm_var(os.fsencode(path), swarm.encode(), name.encode(), n, _ptr(out)), name)
    return out


def write_time(path, dt, time, nproc=1):
    _h5(lib().gs_h5_write_time(os.fsencode(path), float(dt), float(time), int(nproc)), "write_time")


def crand_stream(seed: int, count: int) -> np.ndarray:
    """The C library's rand() after srand(seed) (the generator StrainWeakening's initial damage uses)."""
    out = np.zeros(count, np.int32)
    if lib().gs_crand_stream(int(seed), int(count), _ptr(out)) != 0:
        raise GaleError(E_INVALID, "crand_stream")
    return out


class InitialDamageParams(C.Structure):
    _fields_ = [("fraction", C.c_double), ("factor", C.c_double), ("final_softening_strain", C.c_double),
                ("wavenumber", C.c_double), ("wavenumber_sin_i", C.c_double), ("wavenumber_cos_i", C.c_double),
                ("wavenumber_sin_k", C.c_double), ("wavenumber_cos_k", C.c_double), ("seed", C.c_int64)]


def initial_damage(dim, coord, initialDamageFraction, finalSofteningStrain, randomSeed=0, initialDamageFactor=1.0,
                   initialDamageWavenumber=-1.0, initialDamageWavenumberSinI=-1.0, initialDamageWavenumberCosI=-1.0,
                   initialDamageWavenumberSinK=-1.0, initialDamageWavenumberCosK=-1.0, initialStrainShape=None):
    """StrainWeakening's initial plastic strain of the particles (in layout
    order), XML names; feed it to Problem.set_plastic_strain."""
    c = np.ascontiguousarray(coord, np.float64).reshape(-1, 3)
    d = InitialDamageParams(initialDamageFraction, initialDamageFactor, finalSofteningStrain, initialDamageWavenumber,
                            initialDamageWavenumberSinI, initialDamageWavenumberCosI, initialDamageWavenumberSinK,
                            initialDamageWavenumberCosK, int(randomSeed))
    w = np.zeros(len(c))
    shp = None if initialStrainShape is None else C.byref(initialStrainShape)
    if lib().gs_initial_damage(dim, len(c), _ptr(c), C.byref(d), shp, _ptr(w)) != 0:
        raise GaleError(E_INVALID, "initial_damage")
    return w


def num_nodes(dim, res) -> int:
    return int(np.prod([r + 1 for r in list(res)[:dim]]))


def wall_bc(dim, res, walls, mask=None, values=None):
    """walls: iterable of (wall name, component, value) in WallVC terms."""
    r = np.asarray(_dims(dim, res, 1), np.int32)
    N = num_nodes(dim, res)
    if mask is None:
        mask = np.zeros(N * dim, np.uint8)
    if values is None:
        values = np.zeros(N * dim)
    for wall, comp, val in walls:
        _check(lib().gs_wall_bc(dim, _ptr(r), WALLS[wall], comp, float(val), _ptr(mask), _ptr(values)), "wall_bc")
    return mask, values


def free_slip(dim, res):
    walls = [("left", 0, 0.0), ("right", 0, 0.0), ("bottom", 1, 0.0), ("top", 1, 0.0)]
    if dim == 3:
        walls += [("back", 2, 0.0), ("front", 2, 0.0)]
    return wall_bc(dim, res, walls)


def sphere(material, centre, radius):
    s = Shape()
    s.type, s.material = SHAPE_SPHERE, material
    for i, c in enumerate(list(centre)[:3]):
        s.p[i] = c
    s.p[3] = radius
    return s


def box(material, start, end):
    s = Shape()
    s.type, s.material = SHAPE_BOX, material
    for i in range(min(3, len(start))):
        s.p[i] = start[i]
        s.p[3 + i] = end[i]
    return s


def everywhere(material):
    s = Shape()
    s.type, s.material = SHAPE_EVERYWHERE, material
    return s


def below_sines(material, axis, level, modes=()):
    """modes: (amplitude, k_a, k_b, phase) for level + sum a sin(2 pi (k_a a + k_b b) + phase)."""
    s = Shape()
    s.type, s.material, s.axis = SHAPE_BELOW_SINES, material, axis
    s.p[0] = level
    s.n_modes = len(modes)
    for k, m in enumerate(modes[:8]):
        for j in range(4):
            s.modes[4 * k + j] = m[j]
    return s


def shape_layout(dim, coord, shapes, default=0) -> np.ndarray:
    coord = _f64(coord).reshape(-1, 3)
    mat = np.full(coord.shape[0], default, np.int32)
    arr = (Shape * len(shapes))(*shapes)
    _check(lib().gs_shape_layout(dim, coord.shape[0], _ptr(coord), len(shapes), C.addressof(arr), _ptr(mat)),
           "shape_layout")
    return mat


def gaussian_field(dim, nodes, spacing, corr, seed, vmin, vmax) -> np.ndarray:
    n = np.asarray(_dims(dim, nodes, 1), np.int32)
    sp = np.asarray(_dims(dim, spacing, 1.0))
    out = np.zeros(int(np.prod(n[:dim])))
    _check(lib().gs_gaussian_field(dim, _ptr(n), _ptr(sp), corr, seed, vmin, vmax, _ptr(out)), "gaussian_field")
    return out


# ---------------------------------------------------------------------------
# problem + solver (gale_stokes.h)
# ---------------------------------------------------------------------------
def material(eta0=1.0, density=0.0, yield_type=YIELD_NONE, cohesion=0.0, friction=0.0, min_yield_stress=0.0,
             min_viscosity=0.0, max_strain_rate=0.0) -> dict:
    return dict(eta0=eta0, density=density, yield_type=yield_type, cohesion=cohesion, friction=friction,
                min_yield_stress=min_yield_stress, min_viscosity=min_viscosity, max_strain_rate=max_strain_rate)


def _make_desc(dim, res, lo, hi, coord, material_index, materials, bc_mask=None, bc_value=None, gravity=1.0,
               ghat=(0.0, -1.0, 0.0), eta_clamp=(0.0, 0.0), eta0_particle=None):
    """A gs_problem_desc and the host arrays it points to (keep them alive)."""
    N = num_nodes(dim, res)
    d = ProblemDesc()
    d.dim = dim
    for i, v in enumerate(_dims(dim, res, 1)):
        d.res[i] = v
    for i, v in enumerate(_dims(dim, lo, 0.0)):
        d.lo[i] = v
    for i, v in enumerate(_dims(dim, hi, 1.0)):
        d.hi[i] = v
    for i in range(3):
        d.ghat[i] = ghat[i]
    keep = []
    if bc_mask is not None:
        bc_mask = np.ascontiguousarray(bc_mask, np.uint8).reshape(-1)
        if bc_mask.size != N * dim:
            raise GaleError(E_INVALID, "bc_mask size")
        keep.append(bc_mask)
        d.bc_mask = _ptr(bc_mask)
    if bc_value is not None:
        bc_value = _f64(bc_value).reshape(-1)
        keep.append(bc_value)
        d.bc_value = _ptr(bc_value)
    d.gravity = gravity
    d.eta_clamp_min, d.eta_clamp_max = eta_clamp
    mats = (Material * len(materials))()
    for i, m in enumerate(materials):
        mats[i] = Material(m.get("eta0", 1.0), m.get("density", 0.0), m.get("yield_type", 0), 0,
                           m.get("cohesion", 0.0), m.get("friction", 0.0), m.get("min_yield_stress", 0.0),
                           m.get("min_viscosity", 0.0), m.get("max_strain_rate", 0.0))
    keep.append(mats)
    d.n_materials = len(materials)
    d.materials = C.addressof(mats)
    coord = _f64(coord).reshape(-1, 3)
    mi = np.ascontiguousarray(material_index, np.int32)
    keep += [coord, mi]
    d.n_particles = coord.shape[0]
    d.coord = _ptr(coord)
    d.material = _ptr(mi)
    if eta0_particle is not None:
        e0 = _f64(eta0_particle)
        keep.append(e0)
        d.eta0_particle = _ptr(e0)
    return d, keep


class Problem:
    """FeMesh + MaterialPointsSwarm + IntegrationPointsSwarm on the GPU."""

    def __init__(self, dim, res, lo, hi, coord, material_index, materials, bc_mask=None, bc_value=None,
                 gravity=1.0, ghat=(0.0, -1.0, 0.0), eta_clamp=(0.0, 0.0), eta0_particle=None):
        self.dim = dim
        self.res = tuple(list(res)[:dim])
        self.lo, self.hi = tuple(list(lo)[:dim]), tuple(list(hi)[:dim])
        self.N = num_nodes(dim, res)
        self.E = int(np.prod(self.res))
        d, keep = _make_desc(dim, res, lo, hi, coord, material_index, materials, bc_mask, bc_value, gravity, ghat,
                             eta_clamp, eta0_particle)
        self.P = int(d.n_particles)
        h = C.c_void_p()
        st = lib().gs_problem_create(C.byref(d), C.byref(h))
        self._h = h
        self.create_status = st
        if not h.value:
            raise GaleError(st, "gs_problem_create")
        del keep

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gs_problem_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = np.zeros(5, np.int64)
        _check(lib().gs_problem_info(self._h, _ptr(out)))
        return dict(N=int(out[0]), E=int(out[1]), P=int(out[2]), dim=int(out[3]), escaped=int(out[4]))

    def particles(self):
        coord = np.zeros((self.P, 3))
        owner = np.zeros(self.P, np.int32)
        xi = np.zeros((self.P, self.dim))
        w = np.zeros(self.P)
        _check(lib().gs_problem_particles(self._h, _ptr(coord), _ptr(owner), _ptr(xi), _ptr(w)), "particles")
        return coord, owner, xi, w

    def relocate(self):
        owner = np.zeros(self.P, np.int32)
        st = lib().gs_relocate(self._h, _ptr(owner))
        _check(st, "relocate", allow=(E_ESCAPED,))
        return st, owner

    def set_coords(self, coord):
        c = _f64(coord).reshape(-1, 3)
        st = lib().gs_problem_set_coords(self._h, _ptr(c))
        _check(st, "set_coords", allow=(E_ESCAPED,))
        return st

    def population_control(self, params=None):
        """One PCDVC pass (splits, deletions, Inflow re-population when on),
        then DVC weights; returns (added, deleted) and sets self.rand_draws to
        the rand() calls made.  The particle count (self.P) changes; the
        caller order afterwards is the reference swarm's array order (a
        deleted particle's index taken by the array's last particle, new
        particles appended)."""
        p = params if params is not None else pcdvc_params()
        counts = np.zeros(5, np.int64)
        st = lib().gs_problem_population_control(self._h, C.byref(p), _ptr(counts))
        _check(st, "population_control")
        self.P = self.info()["P"]
        self.rand_draws = int(counts[2])
        self.repopulated_cells, self.repopulation_launches = int(counts[3]), int(counts[4])
        return int(counts[0]), int(counts[1])

    def set_weights(self, kind=WEIGHTS_DVC, resolution=10):
        """The integration swarm's WeightsCalculator: WEIGHTS_CONSTANT
        (ConstantWeights) or WEIGHTS_DVC (DVCWeights, "resolution" or
        resolutionX/Y/Z as a tuple).  The particles are re-mapped at once."""
        r = np.array(_dims(3, [resolution] * 3 if np.isscalar(resolution) else resolution, 10), np.int32)
        st = lib().gs_problem_set_weights(self._h, int(kind), _ptr(r))
        _check(st, "set_weights", allow=(E_ESCAPED,))
        return st

    def set_strain_weakening(self, material, healingRate=0.0, initialSofteningStrain=0.0,
                             finalSofteningStrain=float("inf"), cohesionAfterSoftening=0.0,
                             frictionCoefficientAfterSoftening=0.0):
        """StrainWeakening on `material`'s yield rheology, parameters under their
        XML names (StrainWeakening.cxx:232-290; VonMises / DruckerPrager
        after-softening values).  healingRate=None detaches it."""
        if healingRate is None:
            _check(lib().gs_problem_set_strain_weakening(self._h, int(material), None), "set_strain_weakening")
            return
        sw = np.array([healingRate, initialSofteningStrain, finalSofteningStrain, cohesionAfterSoftening,
                       frictionCoefficientAfterSoftening], np.float64)
        _check(lib().gs_problem_set_strain_weakening(self._h, int(material), _ptr(sw)), "set_strain_weakening")

    def set_strain_limit(self, shape=None, strainLimit=0.0):
        """StrainWeakening "strainLimitedShape" (a sphere(), box() or
        everywhere() shape) and "strainLimit"; shape=None removes it."""
        if shape is None:
            _check(lib().gs_problem_set_strain_limit(self._h, -1, None, 0.0), "set_strain_limit")
            return
        prm = np.array([shape.p[i] for i in range(6)], np.float64)
        _check(lib().gs_problem_set_strain_limit(self._h, int(shape.type), _ptr(prm), float(strainLimit)),
               "set_strain_limit")

    def set_plastic_strain(self, w):
        """Accumulated plastic strain of every particle (caller order)."""
        w = np.ascontiguousarray(w, np.float64)
        if w.shape != (self.P,):
            raise GaleError(E_INVALID, "plastic strain size")
        _check(lib().gs_problem_set_plastic_strain(self._h, _ptr(w)), "set_plastic_strain")

    def plastic_strain(self):
        """(accumulated plastic strain, its rate from the last constitutive pass), caller order."""
        w, inc = np.zeros(self.P), np.zeros(self.P)
        _check(lib().gs_problem_plastic_strain(self._h, _ptr(w), _ptr(inc)), "plastic_strain")
        return w, inc


class PcdvcParams(C.Structure):
    _fields_ = [("upper_t", C.c_double), ("lower_t", C.c_double), ("max_splits", C.c_int32),
                ("max_deletions", C.c_int32), ("split_in_interface_cells", C.c_int32),
                ("delete_in_interface_cells", C.c_int32), ("res", C.c_int32 * 3), ("inflow", C.c_int32),
                ("particles_per_cell", C.c_int32), ("pad_", C.c_int32), ("threshold", C.c_double),
                ("cent_pos_ratio", C.c_double), ("rand_seed", C.c_int64), ("rand_skip", C.c_int64)]


def pcdvc_params(upperT=25.0, lowerT=0.6, maxSplits=3, maxDeletions=2, splitInInterfaceCells=False,
                 deleteInInterfaceCells=False, resolution=10, Inflow=False, particlesPerCell=25, Threshold=0.8,
                 CentPosRatio=0.01, rand_seed=1, rand_skip=0):
    """PCDVC component parameters with the reference's XML names and defaults
    (particlesPerCell is the root dictionary's); rand_seed / rand_skip place
    the Inflow branch on the C library's rand() stream: srand(rand_seed),
    then rand_skip calls already made."""
    r = _dims(3, [resolution] * 3 if np.isscalar(resolution) else resolution, 10)
    return PcdvcParams(upperT, lowerT, maxSplits, maxDeletions, int(bool(splitInInterfaceCells)),
                       int(bool(deleteInInterfaceCells)), (C.c_int32 * 3)(*r), int(bool(Inflow)),
                       int(particlesPerCell), 0, float(Threshold), float(CentPosRatio), int(rand_seed),
                       int(rand_skip))


def dvc_max_particles(dim, resolution=10) -> int:
    r = np.array(_dims(3, [resolution] * 3 if np.isscalar(resolution) else resolution, 10), np.int32)
    return int(lib().gs_dvc_max_particles(int(dim), _ptr(r)))


def dvc_cells(dim, resolution, off, xi):
    """Discrete Voronoi centroids and weights of independent FE cells on the
    device (_DVCWeights_CreateVoronoi + _DVCWeights_GetCentroids per cell):
    cell e holds rows off[e]..off[e+1] of xi (local coordinates)."""
    r = np.array(_dims(3, [resolution] * 3 if np.isscalar(resolution) else resolution, 10), np.int32)
    off = np.ascontiguousarray(off, dtype=np.int32)
    xi = _f64(xi).reshape(-1, dim)
    cent = np.zeros_like(xi)
    w = np.zeros(xi.shape[0])
    _check(lib().gs_dvc_cells(int(dim), _ptr(r), int(len(off) - 1), _ptr(off), _ptr(xi), _ptr(cent), _ptr(w)),
           "gs_dvc_cells")
    return cent, w


def uzawa_params(tolerance=1e-5, maxIterations=1000, minIterations=1, useAbsoluteTolerance=False,
                 Preconditioner=True, precision=PREC_FP64, inner_max_iterations=0, inner_rtol_factor=0.1,
                 inner_rtol_override=0.0, monitor=False, check_every=4, inner_preconditioner=INNER_MULTIGRID,
                 mg_smooth=2, learned=False, flexible_window=16, stagnation_window=50, learned_restarts=5,
                 learned_q_every=0, inner_relax_max=0.0, inner_relax_gain=0.0, mg_smooth_coarse=0) -> UzawaParams:
    """Stokes_SLE_UzawaSolver parameters under their XML names
    (Stokes_SLE_UzawaSolver.cxx:251-257 defaults) plus the inner-solve controls.
    ``learned=True`` selects the learned Schur preconditioner (attach a network
    with ``Stokes_SLE_UzawaSolver.set_learned``).  ``inner_relax_max > 0``
    relaxes the inner tolerance as the outer residual falls (gain
    ``inner_relax_gain``, 0 = 3) and then also selects one smoothing step on
    the multigrid levels below the finest (``mg_smooth_coarse``, 0 = that
    rule; otherwise ``mg_smooth``).  ``check_every`` paces the Jacobi inner
    CG; the multigrid inner CG checks after every iteration."""
    pc = PC_LEARNED if learned else (PC_DIAGONAL if Preconditioner else PC_NONE)
    return UzawaParams(tolerance, maxIterations, minIterations, int(bool(useAbsoluteTolerance)),
                       pc, precision, inner_max_iterations,
                       inner_rtol_factor, inner_rtol_override, int(monitor), check_every,
                       inner_preconditioner, mg_smooth, flexible_window, stagnation_window, learned_restarts,
                       learned_q_every, inner_relax_max, inner_relax_gain, mg_smooth_coarse)


class Stokes_SLE_UzawaSolver:
    """The reference component (Stokes_SLE + Stokes_SLE_UzawaSolver) on the GPU.

    ``Solve`` = _Stokes_SLE_UzawaSolver_Solve, ``GetRhs`` / ``FormResidual``
    the SLE_Solver virtuals, ``NonLinearExecute`` the Picard loop.  Failures
    raise GaleError with the status that mirrors the reference firewall.
    """

    def __init__(self, problem: Problem, params: UzawaParams | None = None, **xml):
        self.problem = problem
        self.params = params or uzawa_params(**xml)
        h = C.c_void_p()
        _check(lib().gs_solver_create(problem._h, C.byref(self.params), C.byref(h)), "gs_solver_create")
        self._h = h
        self.n = problem.N * problem.dim
        self.np = problem.E

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gs_solver_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: UzawaParams):
        self.params = params
        _check(lib().gs_solver_set_params(self._h, C.byref(params)))

    def setup(self, u_prev=None, p_prev=None, yield_on=False):
        u_prev, p_prev = _f64(u_prev), _f64(p_prev)
        _check(lib().gs_solver_setup(self._h, _ptr(u_prev), _ptr(p_prev), int(bool(yield_on))), "setup")

    def Solve(self, raise_on_error=True):
        u = np.zeros(self.n)
        p = np.zeros(self.np)
        rep = Report()
        st = lib().gs_stokes_solve(self._h, _ptr(u), _ptr(p), C.byref(rep))
        if raise_on_error:
            _check(st, "Solve")
        return u, p, rep

    def set_pressure(self, p=None):
        p = _f64(p)
        _check(lib().gs_solver_set_pressure(self._h, _ptr(p)))

    def set_solution(self, u, p):
        """Restart from a previous solution (full nodal velocity, pressure):
        the next setup yields from it, as the reference does when it loads
        from a checkpoint."""
        u, p = _f64(u), _f64(p)
        if u is None or p is None or u.size != self.n or p.size != self.np:
            raise GaleError(E_INVALID, "set_solution: u (N*dim) and p (E) are required")
        _check(lib().gs_solver_set_solution(self._h, _ptr(u), _ptr(p)), "set_solution")

    def NonLinearExecute(self, nonLinearTolerance=0.01, nonLinearMaxIterations=500, nonLinearMinIterations=1,
                         killNonConvergent=False, raise_on_error=True):
        pp = PicardParams(nonLinearTolerance, nonLinearMaxIterations, nonLinearMinIterations,
                          int(bool(killNonConvergent)), 0)
        u = np.zeros(self.n)
        p = np.zeros(self.np)
        rep = Report()
        st = lib().gs_picard_solve(self._h, C.byref(pp), _ptr(u), _ptr(p), C.byref(rep))
        if raise_on_error:
            _check(st, "NonLinearExecute")
        return u, p, rep

    def GetRhs(self):
        out = np.zeros(self.np)
        _check(lib().gs_get_rhs(self._h, _ptr(out)), "GetRhs")
        return out

    def FormResidual(self, p):
        p = _f64(p)
        out = np.zeros(self.np)
        _check(lib().gs_form_residual(self._h, _ptr(p), _ptr(out)), "FormResidual")
        return out

    def compute_dt(self, dtFactor=1.0):
        dt = C.c_double()
        _check(lib().gs_compute_dt(self._h, dtFactor, C.byref(dt)), "compute_dt")
        return dt.value

    def advect_rk2(self, dt):
        st = lib().gs_advect_rk2(self._h, dt)
        _check(st, "advect_rk2", allow=(E_ESCAPED,))
        return st

    def set_learned(self, lpc: "LearnedPreconditioner | None"):
        """Attach the learned Schur preconditioner (used with uzawa_params(learned=True))."""
        self._lpc = lpc
        _check(lib().gs_solver_set_learned(self._h, lpc._h if lpc is not None else None), "set_learned")

    def record_pairs(self, dataset: "PairDataset | None", krylov_every=0):
        """Record training pairs of the following solves into ``dataset``."""
        self._ds = dataset
        _check(lib().gs_solver_record_pairs(self._h, dataset._h if dataset is not None else None,
                                            int(krylov_every)), "record_pairs")

    def particle_fields(self, u=None):
        """StoreStress / StoreViscosity / strain-rate fields of the material
        points (caller order) for the full velocity u (default: the last
        solution)."""
        P, dim = self.problem.P, self.problem.dim
        nst = 3 if dim == 2 else 6
        sr, inv, eta, st = np.zeros((P, nst)), np.zeros(P), np.zeros(P), np.zeros((P, nst))
        uu = None if u is None else _f64(u)
        _check(lib().gs_solver_particle_fields(self._h, _ptr(uu), _ptr(sr), _ptr(inv), _ptr(eta), _ptr(st)),
               "gs_solver_particle_fields")
        return dict(strain_rate=sr, strain_rate_inv=inv, viscosity=eta, stress=st)

    def setup_arrays(self):
        P = self.problem.P
        dg, Q, f, h, eta = np.zeros(self.n), np.zeros(self.np), np.zeros(self.n), np.zeros(self.np), np.zeros(P)
        _check(lib().gs_solver_setup_arrays(self._h, _ptr(dg), _ptr(Q), _ptr(f), _ptr(h), _ptr(eta)))
        return dict(diagK=dg, Q=Q, f=f, h=h, eta=eta)

    # device-pointer operator applies (torch tensors on cuda)
    def apply_K(self, x, y):
        _check(lib().gs_apply_K(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())), "apply_K")

    def apply_K_noalloc(self, x, y):
        """gs_apply_K without any allocation (static per-device scratch)."""
        _check(lib().gs_apply_K_noalloc(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())),
               "apply_K_noalloc")

    def apply_K_fp32(self, x, y):
        _check(lib().gs_apply_K_fp32(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())), "apply_K_fp32")

    def apply_K_moments(self, x, y):
        _check(lib().gs_apply_K_moments(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr())),
               "apply_K_moments")

    def apply_G(self, p, y):
        _check(lib().gs_apply_G(self._h, C.c_void_p(p.data_ptr()), C.c_void_p(y.data_ptr())), "apply_G")

    def apply_Gt(self, x, q):
        _check(lib().gs_apply_Gt(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(q.data_ptr())), "apply_Gt")

    def sync(self):
        _check(lib().gs_solver_sync(self._h))

    def time_apply_K(self, form=APPLY_PARTICLES_FP64, n_rep=20):
        ms = C.c_double()
        _check(lib().gs_time_apply_K(self._h, form, n_rep, C.byref(ms)), "time_apply_K")
        return ms.value

    def time_apply(self, form=APPLY_MOMENTS_FP64, n_rep=20, cold=True):
        """Average device ms per apply; cold=True cycles through operand copies
        larger than the last-level cache so each apply streams from HBM."""
        ms = C.c_double()
        _check(lib().gs_time_apply(self._h, form, n_rep, 1 if cold else 0, C.byref(ms)), "time_apply")
        return ms.value


# ---------------------------------------------------------------------------
# learned Schur-complement preconditioner (include/gale_learned.h)
# ---------------------------------------------------------------------------
class TrainParams(C.Structure):
    _fields_ = [
        ("steps", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("lr_final", C.c_float),
        ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("clip", C.c_float),
        ("warmup", C.c_int32), ("first", C.c_int32), ("count", C.c_int32), ("seed", C.c_uint32),
    ]


def _dptr(t):
    """Device pointer of a torch CUDA tensor (float32, contiguous)."""
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "expected a contiguous device tensor"
    return C.c_void_p(t.data_ptr())


class PairDataset:
    """Training pairs (S e = r) recorded on the device by plain Uzawa solves."""

    def __init__(self, res, capacity):
        nx, ny, nz = (res, res, res) if np.isscalar(res) else tuple(res)
        self.shape = (nx, ny, nz)
        self.V = nx * ny * nz
        h = C.c_void_p()
        _check(lib().gs_lpc_dataset_create(nx, ny, nz, int(capacity), C.byref(h)), "dataset_create")
        self._h = h

    def __len__(self):
        return int(lib().gs_lpc_dataset_size(self._h))

    def get(self, idx):
        inp = np.zeros(2 * self.V, np.float32)
        tgt = np.zeros(self.V, np.float32)
        w = C.c_float()
        pj = C.c_int32()
        _check(lib().gs_lpc_dataset_get(self._h, int(idx), _ptr(inp), _ptr(tgt), C.byref(w), C.byref(pj)))
        return inp.reshape(self.V, 2), tgt, w.value, pj.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gs_lpc_dataset_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LearnedPreconditioner:
    """BF16 matrix-core conv encoder-decoder preconditioning the Schur residual."""

    def __init__(self, seed=1):
        h = C.c_void_p()
        _check(lib().gs_lpc_create(C.c_uint64(seed), C.byref(h)), "lpc_create")
        self._h = h

    @staticmethod
    def num_params():
        return int(lib().gs_lpc_num_params())

    def get_params(self):
        out = np.zeros(self.num_params(), np.float32)
        _check(lib().gs_lpc_get_params(self._h, _ptr(out)))
        return out

    def set_params(self, params):
        params = np.ascontiguousarray(params, np.float32)
        assert params.size == self.num_params()
        _check(lib().gs_lpc_set_params(self._h, _ptr(params)))

    def forward(self, d_in, d_out):
        """d_in: float32 device tensor [B, nz, ny, nx, 2]; d_out: [B, nz, ny, nx]."""
        B, nz, ny, nx, _ = d_in.shape
        _check(lib().gs_lpc_forward(self._h, nx, ny, nz, B, _dptr(d_in), _dptr(d_out)), "lpc_forward")

    def backward(self, d_in, d_gout):
        B, nz, ny, nx, _ = d_in.shape
        g = np.zeros(self.num_params(), np.float32)
        _check(lib().gs_lpc_backward(self._h, nx, ny, nz, B, _dptr(d_in), _dptr(d_gout), _ptr(g)), "lpc_backward")
        return g

    def loss(self, d_out, d_target, weights, proj, d_gout):
        B, nz, ny, nx = d_out.shape
        w = np.ascontiguousarray(weights, np.float32)
        pj = np.ascontiguousarray(proj, np.int32)
        L = C.c_double()
        _check(lib().gs_lpc_loss(self._h, nx, ny, nz, B, _dptr(d_out), _dptr(d_target), _ptr(w), _ptr(pj),
                                 _dptr(d_gout), C.byref(L)), "lpc_loss")
        return L.value

    def train(self, dataset: PairDataset, steps, batch=8, lr=1e-3, lr_final=1e-5, clip=1.0, warmup=50,
              first=0, count=0, seed=1, beta1=0.9, beta2=0.999, eps=1e-8):
        p = TrainParams(int(steps), int(batch), lr, lr_final, beta1, beta2, eps, clip, int(warmup), int(first),
                        int(count), int(seed))
        hist = np.zeros(max(int(steps), 1), np.float32)
        _check(lib().gs_lpc_train(self._h, dataset._h, C.byref(p), _ptr(hist)), "lpc_train")
        return hist[:steps]

    def evaluate(self, dataset: PairDataset, first, count):
        L = C.c_double()
        _check(lib().gs_lpc_eval(self._h, dataset._h, int(first), int(count), C.byref(L)), "lpc_eval")
        return L.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gs_lpc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lpc_selftest():
    e = C.c_float()
    _check(lib().gs_lpc_selftest(C.byref(e)), "lpc_selftest")
    return e.value


# ---------------------------------------------------------------------------
# batched solve of many small 2D problems (include/gale_batch.h)
# ---------------------------------------------------------------------------
class BatchParams(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double), ("max_iterations", C.c_int32), ("max_inner", C.c_int32),
        ("inner_rtol_factor", C.c_double), ("mg_smooth", C.c_int32), ("use_graph", C.c_int32),
        ("inner_relax_max", C.c_double), ("flexible_window", C.c_int32), ("pad_", C.c_int32),
    ]


class BatchResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("uzawa_iterations", C.c_int32), ("inner_iterations", C.c_int64),
        ("rel_residual", C.c_double), ("seconds", C.c_double),
    ]


TILE_CANDIDATES = [(256, 16), (256, 32), (512, 32), (512, 64), (1024, 32), (1024, 64)]


def batch_params(tolerance=1e-6, maxIterations=5000, max_inner=20000, inner_rtol_factor=0.1, mg_smooth=2,
                 use_graph=False, inner_relax_max=0.0, flexible_window=8) -> BatchParams:
    """inner_relax_max > 0 selects the inexact (relaxed) Uzawa loop with a
    flexible outer iteration and a true-residual check; 0 the reference loop."""
    return BatchParams(tolerance, maxIterations, max_inner, inner_rtol_factor, mg_smooth, int(bool(use_graph)),
                       inner_relax_max, flexible_window, 0)


class BatchSolver:
    """Many independent 2D problems solved in one persistent launch."""

    def __init__(self, problems):
        """problems: list of Problem keyword dicts (Config.problem_kwargs())."""
        self.count = len(problems)
        descs = (ProblemDesc * self.count)()
        keep = []
        self.shapes = []
        for i, kw in enumerate(problems):
            d, k = _make_desc(**kw)
            descs[i] = d
            keep.append(k)
            res = tuple(kw["res"])[:2]
            self.shapes.append(res)
        h = C.c_void_p()
        _check(lib().gs_batch_create(self.count, descs, C.byref(h)), "gs_batch_create")
        self._h = h
        del keep

    def set_tile(self, block, tx):
        _check(lib().gs_batch_set_tile(self._h, int(block), int(tx)), "gs_batch_set_tile")

    def autotune(self, params=None):
        params = params or batch_params()
        b, t = C.c_int32(), C.c_int32()
        ms = np.zeros(len(TILE_CANDIDATES), np.float32)
        _check(lib().gs_batch_autotune(self._h, C.byref(params), C.byref(b), C.byref(t), _ptr(ms)), "autotune")
        return (b.value, t.value), dict(zip(TILE_CANDIDATES, map(float, ms)))

    def solve(self, params=None, raise_on_error=True):
        params = params or batch_params()
        res = (BatchResult * self.count)()
        sec = C.c_double()
        st = lib().gs_batch_solve(self._h, C.byref(params), res, C.byref(sec))
        if raise_on_error:
            _check(st, "gs_batch_solve")
        return st, [dict(status=r.status, uzawa_iterations=r.uzawa_iterations,
                         inner_iterations=r.inner_iterations, rel_residual=r.rel_residual, seconds=r.seconds)
                    for r in res], sec.value

    def solution(self, i):
        nx, ny = self.shapes[i]
        u = np.zeros(2 * (nx + 1) * (ny + 1))
        p = np.zeros(nx * ny)
        _check(lib().gs_batch_solution(self._h, int(i), _ptr(u), _ptr(p)), "gs_batch_solution")
        return u, p

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().gs_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

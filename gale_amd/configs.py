"""Synthetic Q1-P0 PIC Stokes workloads (the XML inputs of the reference,
restated as builders).  Every builder is deterministic for a given seed.

sinker_2d              unit square, free slip, dense viscous disc (RayTay_Example-like)
lithosphere_2d         4x1 box, Drucker-Prager lid, extension +-1 on the sides
rayleigh_taylor_3d     unit cube, dense layer over light layer, sine-perturbed interface
random_viscosity_3d    unit cube, log10 eta Gaussian random field in [-3, 3], plastic lid, dense sphere
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import (YIELD_DRUCKER_PRAGER, below_sines, box, everywhere, free_slip, gaussian_field, layout_random,
               material, num_nodes, shape_layout, sphere, wall_bc)


@dataclass
class Config:
    name: str
    dim: int
    res: tuple
    lo: tuple
    hi: tuple
    coord: np.ndarray
    material: np.ndarray
    materials: list
    bc_mask: np.ndarray
    bc_value: np.ndarray
    gravity: float = 1.0
    ghat: tuple = (0.0, -1.0, 0.0)
    eta_clamp: tuple = (0.0, 0.0)
    eta0_particle: np.ndarray | None = None
    tolerance: float = 1e-6
    picard: dict = field(default_factory=dict)
    steps: int = 0

    @property
    def E(self):
        return int(np.prod(self.res))

    @property
    def N(self):
        return num_nodes(self.dim, self.res)

    @property
    def P(self):
        return int(self.coord.shape[0])

    def problem_kwargs(self):
        return dict(dim=self.dim, res=self.res, lo=self.lo, hi=self.hi, coord=self.coord,
                    material_index=self.material, materials=self.materials, bc_mask=self.bc_mask,
                    bc_value=self.bc_value, gravity=self.gravity, ghat=self.ghat, eta_clamp=self.eta_clamp,
                    eta0_particle=self.eta0_particle)

    def oracle_kwargs(self):
        kw = self.problem_kwargs()
        kw["material"] = kw.pop("material_index")
        return kw


def sinker_2d(res=32, ppc=20, seed=13, eta_disc=1e3, rho_disc=1.1) -> Config:
    """2D sinker: unit square, free-slip walls, background rho 1 eta 1, disc
    r 0.1 at (0.5, 0.6) with rho 1.1 and eta 1e3, gravity 1 along -y."""
    dim, r = 2, (res, res)
    lo, hi = (0.0, 0.0), (1.0, 1.0)
    coord = layout_random(dim, r, lo, hi, ppc, seed)
    mats = [material(eta0=1.0, density=1.0), material(eta0=eta_disc, density=rho_disc)]
    mat = shape_layout(dim, coord, [everywhere(0), sphere(1, (0.5, 0.6, 0.0), 0.1)])
    mask, vals = free_slip(dim, r)
    return Config("sinker_2d", dim, r, lo, hi, coord, mat, mats, mask, vals, steps=1)


def sinker_instance(rng, ppc=20, res_range=(16, 128), seed=1):
    """One instance of the batched sinker family: square mesh n x n with n
    drawn uniformly from the multiples of 8 in res_range (so that every
    instance has a nested multigrid hierarchy down to <= 16^2), 1-5 discs (radius in [0.05, 0.15], density contrast
    in [0.05, 0.2] over the background rho 1, log10 viscosity contrast
    uniform in [-2, 4] over the background eta 1), free-slip unit square,
    gravity 1 along -y.  rng: numpy Generator drawing the instance parameters;
    seed: the particle layout seed."""
    lo8, hi8 = -(-res_range[0] // 8), res_range[1] // 8
    n = 8 * int(rng.integers(lo8, hi8 + 1)) if hi8 >= lo8 else int(res_range[0])
    dim, r = 2, (n, n)
    lo, hi = (0.0, 0.0), (1.0, 1.0)
    coord = layout_random(dim, r, lo, hi, ppc, seed)
    k = int(rng.integers(1, 6))
    mats = [material(eta0=1.0, density=1.0)]
    shapes = [everywhere(0)]
    for i in range(k):
        rad = float(rng.uniform(0.05, 0.15))
        cx, cy = (float(v) for v in rng.uniform(rad, 1.0 - rad, 2))
        drho = float(rng.uniform(0.05, 0.2))
        leta = float(rng.uniform(-2.0, 4.0))
        mats.append(material(eta0=10.0 ** leta, density=1.0 + drho))
        shapes.append(sphere(i + 1, (cx, cy, 0.0), rad))
    mat = shape_layout(dim, coord, shapes)
    mask, vals = free_slip(dim, r)
    return Config(f"sinker_instance_{n}", dim, r, lo, hi, coord, mat, mats, mask, vals, steps=1)


def sinker_batch(count=256, seed=2024, ppc=20, res_range=(16, 128)):
    """The batched workload: count independent sinker instances of uneven mesh
    size (see sinker_instance), reproducible from seed."""
    rng = np.random.default_rng(seed)
    return [sinker_instance(rng, ppc=ppc, res_range=res_range, seed=seed * 1000 + i) for i in range(count)]


def lithosphere_2d(nx=256, ny=64, ppc=40, seed=13) -> Config:
    """2D extension of a plastic lid: 4x1 box, lid y >= 0.75 Drucker-Prager
    (cohesion 1e-2, friction angle 30 deg, eta 1e2, minimum viscosity 1e-3),
    mantle eta 1, weak seed 0.05 square of eta 1e-2 at (2.0, 0.875); vx -1 / +1
    on the left / right walls, vy 0 on the base, free top; rho 0 (kinematic
    forcing only); viscosity clamp [1e-3, 1e3]; Picard <= 20 to 1e-4."""
    dim, r = 2, (nx, ny)
    lo, hi = (0.0, 0.0), (4.0, 1.0)
    coord = layout_random(dim, r, lo, hi, ppc, seed)
    lid = material(eta0=1e2, density=0.0, yield_type=YIELD_DRUCKER_PRAGER, cohesion=1e-2,
                   friction=math.tan(math.radians(30.0)), min_viscosity=1e-3)
    mats = [material(eta0=1.0), lid, material(eta0=1e-2)]
    mat = shape_layout(dim, coord, [everywhere(0), box(1, (0.0, 0.75, 0.0), (4.0, 1.0, 0.0)),
                                    box(2, (1.975, 0.85, 0.0), (2.025, 0.9, 0.0))])
    mask, vals = wall_bc(dim, r, [("left", 0, -1.0), ("right", 0, 1.0), ("bottom", 1, 0.0)])
    return Config("lithosphere_2d", dim, r, lo, hi, coord, mat, mats, mask, vals, gravity=0.0,
                  eta_clamp=(1e-3, 1e3), picard=dict(nonLinearTolerance=1e-4, nonLinearMaxIterations=20))


def lithosphere_3d(nx=16, ny=4, nz=4, ppc=10, seed=13) -> Config:
    """3D counterpart of lithosphere_2d: a 4 x 1 x 1 box, lid y >= 0.75
    Drucker-Prager (cohesion 1e-2, friction angle 30 deg, eta 1e2, minimum
    viscosity 1e-3), mantle eta 1, a weak 0.05 x 0.05 x 1 seed of eta 1e-2 at
    x = 2.0, y = 0.875 through the depth; vx -1 / +1 on the left / right
    walls, vy 0 on the base, vz 0 on the back and front walls, free top;
    rho 0; viscosity clamp [1e-3, 1e3]; Picard <= 20 to 1e-4."""
    dim, r = 3, (nx, ny, nz)
    lo, hi = (0.0, 0.0, 0.0), (4.0, 1.0, 1.0)
    coord = layout_random(dim, r, lo, hi, ppc, seed)
    lid = material(eta0=1e2, density=0.0, yield_type=YIELD_DRUCKER_PRAGER, cohesion=1e-2,
                   friction=math.tan(math.radians(30.0)), min_viscosity=1e-3)
    mats = [material(eta0=1.0), lid, material(eta0=1e-2)]
    mat = shape_layout(dim, coord, [everywhere(0), box(1, (0.0, 0.75, 0.0), (4.0, 1.0, 1.0)),
                                    box(2, (1.975, 0.85, 0.0), (2.025, 0.9, 1.0))])
    mask, vals = wall_bc(dim, r, [("left", 0, -1.0), ("right", 0, 1.0), ("bottom", 1, 0.0), ("back", 2, 0.0),
                                  ("front", 2, 0.0)])
    return Config("lithosphere_3d", dim, r, lo, hi, coord, mat, mats, mask, vals, gravity=0.0,
                  eta_clamp=(1e-3, 1e3), picard=dict(nonLinearTolerance=1e-4, nonLinearMaxIterations=20))


def rayleigh_taylor_3d(res=64, ppc=20, seed=13, mode_seed=1003, steps=10) -> Config:
    """3D Rayleigh-Taylor: unit cube, dense layer (rho 1, eta 1e2) over light
    (rho 0, eta 1); interface at y = 0.5 + four seeded sine modes whose
    amplitudes sum to 0.02; free slip; gravity -y."""
    dim, r = 3, (res, res, res)
    lo, hi = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    coord = layout_random(dim, r, lo, hi, ppc, seed)
    rng = np.random.default_rng(mode_seed)
    amps = rng.uniform(0.5, 1.0, 4)
    amps *= 0.02 / amps.sum()
    modes = [(amps[k], 0.5 * (k + 1), 0.5 * (4 - k), rng.uniform(0, 2 * math.pi)) for k in range(4)]
    mats = [material(eta0=1e2, density=1.0), material(eta0=1.0, density=0.0)]
    mat = shape_layout(dim, coord, [everywhere(0), below_sines(1, 1, 0.5, modes)])
    mask, vals = free_slip(dim, r)
    return Config("rayleigh_taylor_3d", dim, r, lo, hi, coord, mat, mats, mask, vals, steps=steps)


def random_viscosity_3d(res=96, ppc=16, seed=13, field_seed=1005, corr=0.1) -> Config:
    """3D viscosity-contrast solve: unit cube, log10 eta a seeded Gaussian
    random field (correlation length 0.1) min-max scaled to [-3, 3] and
    sampled trilinearly at the particles, a Drucker-Prager lid of the
    lithosphere type on top (y >= 0.75, eta 1e2), and a dense sphere (rho 1,
    r 0.15 at the centre) in a rho 0 background driving the flow; free slip."""
    dim, r = 3, (res, res, res)
    lo, hi = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    coord = layout_random(dim, r, lo, hi, ppc, seed)
    nodes = (res + 1,) * 3
    h = 1.0 / res
    fld = gaussian_field(dim, nodes, (h, h, h), corr, field_seed, -3.0, 3.0).reshape(nodes[::-1])  # [k][j][i]
    # trilinear interpolation of log10 eta at the particles
    t = coord * res
    i0 = np.minimum(np.floor(t).astype(np.int64), res - 1)
    fr = t - i0
    le = np.zeros(coord.shape[0])
    for a in range(8):
        ax, ay, az = a & 1, (a >> 1) & 1, (a >> 2) & 1
        wgt = ((fr[:, 0] if ax else 1 - fr[:, 0]) * (fr[:, 1] if ay else 1 - fr[:, 1]) *
               (fr[:, 2] if az else 1 - fr[:, 2]))
        le += wgt * fld[i0[:, 2] + az, i0[:, 1] + ay, i0[:, 0] + ax]
    eta0p = 10.0 ** le
    lid = material(eta0=1e2, density=0.0, yield_type=YIELD_DRUCKER_PRAGER, cohesion=1e-2,
                   friction=math.tan(math.radians(30.0)), min_viscosity=1e-3)
    mats = [material(eta0=1.0, density=0.0), lid, material(eta0=1.0, density=1.0)]
    mat = shape_layout(dim, coord, [everywhere(0), box(1, (0.0, 0.75, 0.0), (1.0, 1.0, 1.0)),
                                    sphere(2, (0.5, 0.5, 0.5), 0.15)])
    eta0p[mat == 1] = 1e2
    mask, vals = free_slip(dim, r)
    return Config("random_viscosity_3d", dim, r, lo, hi, coord, mat, mats, mask, vals, eta0_particle=eta0p,
                  eta_clamp=(1e-3, 1e3), picard=dict(nonLinearTolerance=1e-4, nonLinearMaxIterations=20))


BUILDERS = {
    "sinker_2d": sinker_2d,
    "lithosphere_2d": lithosphere_2d,
    "lithosphere_3d": lithosphere_3d,
    "rayleigh_taylor_3d": rayleigh_taylor_3d,
    "random_viscosity_3d": random_viscosity_3d,
}

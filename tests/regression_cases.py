"""The reference's regression fixtures at step 10, as Stokes problems.

tests/golden/<case>_step10/ holds the HDF5 checkpoints (Mesh, VelocityField,
PressureField, materialSwarm, timeInfo) that the reference's own regression
suites compare against (Underworld/SysTest/RegressionTests/expected/
RayleighTaylorBenchmark.xml, PICellerator/SysTest/RegressionTests/expected/
LidDrivenPIC2D.xml, StgFEM/SysTest/RegressionTests/expected/Multigrid.xml).
The StgFEM case integrates at 2 x 2 Gauss points (no swarm, nothing
advected).  For the PIC cases the checkpoint is written after the step-10 solve and
the particle advection that follows it, so the swarm it holds is one RK2 step
(TimeIntegrand.cxx:265-335) past the one that step's solve integrated over;
load() undoes that step with the saved velocity field (fixed-point iteration
on x1 = x0 + dt v(x0 + dt/2 v(x0)), bilinear interpolation of the nodal
field) and returns the problem whose Stokes solution the saved velocity and
pressure fields are.  Particles that PCDVC population control added or
removed after the advection are the residual difference."""
import os

import numpy as np

from conftest import GOLDEN

# RayleighTaylorBenchmark.xml (Underworld/InputFiles): 64 x 64 over 0.9142 x 1,
# background rho 1 eta 1, light layer rho 0 eta 1, gravity 1; BCs from
# VariableConditions/velocityBCs.freeslipSides.xml (no slip top and bottom,
# vx = 0 on the sides); weights PCDVC with a 10 x 10 sub-grid
# (PICellerator/Apps/PIC_Components/MaterialPointSwarm.xml).
# LidDrivenPIC2D.xml (PICellerator/SysTest/RegressionTests): 20 x 20 unit
# square, isoviscous eta 1 (PIC_Components/StokesFlowUzawa.xml), no buoyancy,
# lidDrivenVelocityBCs.xml: vy = 0 bottom and top, vx = 0 sides, vx on the top
# from StgFEM_StandardConditionFunctions_PartialLid_TopLayer
# (StgFEM/plugins/StandardConditionFunctions/StandardConditionFunctions.cxx:547-567).
# Multigrid.xml (StgFEM/SysTest/RegressionTests): 32 x 16 over 2 x 1, Gauss
# swarm 2 x 2 (StgFEM/Apps/StgFEM_Components/GaussSwarm.xml), isoviscous, vx = 0
# on the sides, vy = 0 top and bottom, vx = 1 on the whole top (the top wall is
# the last entry of the CompositeVC, so it sets the corners).
CASES = {
    "rayleigh_taylor": dict(folder="rayleigh_taylor_step10", res=(64, 64), lo=(0.0, 0.0), hi=(0.9142, 1.0),
                            gravity=1.0, materials=[(1.0, 1.0), (1.0, 0.0)],
                            walls=[("bottom", 0, 0.0), ("bottom", 1, 0.0), ("top", 0, 0.0), ("top", 1, 0.0),
                                   ("left", 0, 0.0), ("right", 0, 0.0)], lid=False, gauss=False, mesh="Mesh"),
    "lid_driven": dict(folder="lid_driven_pic_step10", res=(20, 20), lo=(0.0, 0.0), hi=(1.0, 1.0),
                       gravity=0.0, materials=[(1.0, 0.0)],
                       walls=[("bottom", 1, 0.0), ("left", 0, 0.0), ("right", 0, 0.0), ("top", 1, 0.0)], lid=True,
                       gauss=False, mesh="Mesh"),
    # the regression suite's own low-resolution run (lowres_np1.dat: 10 x 10), checkpoint at step 5
    "rayleigh_taylor_lowres": dict(folder="rayleigh_taylor_lowres_step5", res=(10, 10), lo=(0.0, 0.0),
                                   hi=(0.9142, 1.0), gravity=1.0, materials=[(1.0, 1.0), (1.0, 0.0)],
                                   walls=[("bottom", 0, 0.0), ("bottom", 1, 0.0), ("top", 0, 0.0), ("top", 1, 0.0),
                                          ("left", 0, 0.0), ("right", 0, 0.0)], lid=False, gauss=False,
                                   mesh="Mesh.linearMesh", step="00005"),
    "stokes_multigrid": dict(folder="stokes_multigrid_step10", res=(32, 16), lo=(0.0, 0.0), hi=(2.0, 1.0),
                             gravity=0.0, materials=[(1.0, 0.0)],
                             walls=[("left", 0, 0.0), ("right", 0, 0.0), ("bottom", 1, 0.0), ("top", 1, 0.0),
                                    ("top", 0, 1.0)], lid=False, gauss=True, mesh="Mesh.linearMesh"),
}
PIC = [k for k, c in CASES.items() if not c["gauss"]]


def path(case, what):
    return os.path.join(GOLDEN, CASES[case]["folder"], what + "." + CASES[case].get("step", "00010") + ".h5")


def _interp(U, pts, res, lo, hi):
    h = (np.asarray(hi) - np.asarray(lo)) / np.asarray(res)
    q = (pts[:, :2] - np.asarray(lo)) / h
    i = np.minimum(np.maximum(np.floor(q).astype(np.int64), 0), np.asarray(res) - 1)
    f = q - i
    nx = res[0] + 1
    n = lambda a, b: (i[:, 1] + b) * nx + i[:, 0] + a
    return (((1 - f[:, 0]) * (1 - f[:, 1]))[:, None] * U[n(0, 0)] + (f[:, 0] * (1 - f[:, 1]))[:, None] * U[n(1, 0)]
            + ((1 - f[:, 0]) * f[:, 1])[:, None] * U[n(0, 1)] + (f[:, 0] * f[:, 1])[:, None] * U[n(1, 1)])


def load(case, gale):
    """coord (P, 3) before the last advection, material (P,), the saved
    velocity (N, 2) and P0 pressure (E,), dt, bc mask and values."""
    c = CASES[case]
    res, lo, hi = c["res"], c["lo"], c["hi"]
    N = (res[0] + 1) * (res[1] + 1)
    E = res[0] * res[1]
    u = gale.read_field(path(case, "VelocityField"), 2, 2, N)
    p = gale.read_field(path(case, "PressureField"), 2, 1, E)[:, 0]
    dt = float(gale.h5_read(path(case, "timeInfo"), "Dt")[0])
    if c["gauss"]:
        coord = gale.layout_gauss(2, res, lo, hi)
        mat = np.zeros(len(coord), np.int32)
    else:
        x1, mat = gale.read_swarm(path(case, "materialSwarm"), "materialSwarm", 2)
        x0 = x1[:, :2].copy()
        for _ in range(40):
            mid = x0 + 0.5 * dt * _interp(u, x0, res, lo, hi)
            x0 = x1[:, :2] - dt * _interp(u, mid, res, lo, hi)
        coord = np.zeros_like(x1)
        coord[:, :2] = x0
    mask, vals = gale.wall_bc(2, res, c["walls"])
    if c["lid"]:
        # PartialLid_TopLayer: vx = 1 where min + margin < x < max - margin,
        # margin 1.1 x the minimum node separation
        margin = 1.1 * min((hi[d] - lo[d]) / res[d] for d in range(2))
        for i in range(res[0] + 1):
            x = lo[0] + (i / res[0]) * (hi[0] - lo[0])
            node = res[1] * (res[0] + 1) + i
            mask[2 * node] = 1
            vals[2 * node] = 1.0 if lo[0] + margin < x < hi[0] - margin else 0.0
    return dict(coord=coord, material=mat, u=u, p=p, dt=dt, bc_mask=mask, bc_value=vals)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))

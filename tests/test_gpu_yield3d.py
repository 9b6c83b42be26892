"""3D viscoplastic solves on the GPU against the oracle: the 3D Drucker-Prager
branch (factor 2 cos(phi) / (sqrt(3) (3 - sin(phi))), cohesion x 3,
DruckerPrager.cxx:376-405) yielding inside a Picard sequence, and the rule
that decides whether a first solve yields: previousSolutionExists, set by the
first nonlinear execute and kept for later time steps or set by a restart
(ConstitutiveMatrixCartesian.cxx:270-273, ConstitutiveMatrix.cxx:127-130,
YieldRheology.cxx:203-204)."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def _lith3d():
    from gale_amd import configs
    return configs.lithosphere_3d(nx=8, ny=4, nz=4, ppc=6)


def _yield_sets(gale, oracle, g, o, s, u, p, cfg):
    """Particle viscosities of one yielding constitutive pass from the same
    (u, p), GPU and oracle, in caller order."""
    s.setup(u, p, yield_on=True)
    eta_g = s.setup_arrays()["eta"]
    st, eta_o, yl_o = o.viscosity(u, p, 1)
    assert st == 0
    return eta_g, eta_o, yl_o


def test_picard_drucker_prager_3d_parity(gale, oracle):
    from test_gpu_parity import otight, pair, tight
    cfg = _lith3d()
    g, o = pair(gale, oracle, cfg)
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale, tolerance=1e-11))
    u, p, rep = s.NonLinearExecute(nonLinearTolerance=1e-4, nonLinearMaxIterations=20)
    st, ou, op, orep = o.picard(otight(oracle, tolerance=1e-11), tolerance=1e-4, max_iterations=20)
    print(f"3D Picard: GPU {rep.picard_iterations} its, oracle {orep.picard_iterations}", flush=True)
    assert st == 0 and rep.status == 0
    assert rep.picard_iterations == orep.picard_iterations > 2
    assert rel(u, ou) < 1e-8
    assert rel(p, op) < 1e-8
    # one yielding constitutive pass from the same solution: equal viscosities
    # and the same set of yielded points, and the lid yields
    eta_g, eta_o, yl_o = _yield_sets(gale, oracle, g, o, s, u, p, cfg)
    lid = cfg.material == 1
    assert yl_o[lid].sum() > 0.5 * lid.sum()
    assert rel(eta_g, eta_o) < 1e-12
    np.testing.assert_array_equal(eta_g[lid] < 1e2, yl_o[lid] == 1)
    # the GPU's yielded viscosity against the 3D closed form, independent of
    # the oracle: eta = max(eta_min, tau_y / (2 eps_II)) with
    # tau_y = max(mu f p_e + 3 C f, C f), f = 2 cos(phi) / (sqrt(3) (3 - sin(phi))),
    # eps_II the second invariant of the deviatoric strain rate at the point
    # (the trace removed, VonMises.cxx:219-240)
    m = cfg.materials[1]
    mu, C = m["friction"], m["cohesion"]
    cphi = 1.0 / np.sqrt(1 + mu * mu)
    f = 2 * cphi / (np.sqrt(3.0) * (3 - mu * cphi))
    owner = g.particles()[1]
    sr = s.particle_fields(u)["strain_rate"]          # xx, yy, zz, xy, xz, yz
    dev = sr[:, :3] - sr[:, :3].sum(1, keepdims=True) / 3.0
    inv = np.sqrt(0.5 * (dev ** 2).sum(1) + (sr[:, 3:] ** 2).sum(1))
    y = lid & (yl_o == 1)
    tau_y = np.maximum(mu * f * p[owner[y]] + 3 * C * f, C * f)
    expect = np.maximum(m["min_viscosity"], tau_y / (2 * inv[y]))
    np.testing.assert_allclose(eta_g[y], expect, rtol=1e-10)
    s.close()
    g.close()


def test_second_time_step_first_solve_yields(gale, oracle):
    """After a nonlinear execute, the next time step's first solve yields from
    the previous solution: the Picard sequence of step 2 matches the oracle's
    started from the previous step (same iteration count, solution to 1e-8),
    and a one-iteration Picard of a continuing solver yields where a fresh
    solver on the same particles does not."""
    from test_gpu_parity import otight, pair, tight
    cfg = _lith3d()
    g, o = pair(gale, oracle, cfg)
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale, tolerance=1e-11))
    u, p, rep = s.NonLinearExecute(nonLinearTolerance=1e-4, nonLinearMaxIterations=20)
    st, ou, op, orep = o.picard(otight(oracle, tolerance=1e-11), tolerance=1e-4, max_iterations=20)
    assert st == 0 and rep.status == 0
    # the side walls pull outwards at |vx| = 1: keep every particle inside
    x = g.particles()[0][:, 0]
    dt = min(0.5 * s.compute_dt(0.5), 0.25 * min(x.min() - cfg.lo[0], cfg.hi[0] - x.max()))
    assert s.advect_rk2(dt) == 0
    o.set_coords(g.particles()[0])
    u2, p2, rep2 = s.NonLinearExecute(nonLinearTolerance=1e-4, nonLinearMaxIterations=20)
    st, ou2, op2, orep2 = o.picard(otight(oracle, tolerance=1e-11), tolerance=1e-4, max_iterations=20,
                                   u_prev=ou, p_prev=op)
    print(f"step 2 Picard: GPU {rep2.picard_iterations}, oracle {orep2.picard_iterations}", flush=True)
    assert st == 0 and rep2.status == 0
    assert rep2.picard_iterations == orep2.picard_iterations
    assert rel(u2, ou2) < 1e-8 and rel(p2, op2) < 1e-8
    # the first solve of a later step alone: yielding on (continuing solver)
    # against off (a fresh solver on the same particles)
    s.NonLinearExecute(nonLinearTolerance=1e-4, nonLinearMaxIterations=1)
    eta_cont = s.setup_arrays()["eta"]
    fresh = gale.Stokes_SLE_UzawaSolver(g, tight(gale, tolerance=1e-11))
    fresh.NonLinearExecute(nonLinearTolerance=1e-4, nonLinearMaxIterations=1)
    eta_fresh = fresh.setup_arrays()["eta"]
    lid = cfg.material == 1
    assert np.all(eta_fresh[lid] == 1e2)
    assert np.sum(eta_cont[lid] < 1e2) > 0.5 * lid.sum()
    fresh.close()
    s.close()
    g.close()


def test_restart_yields_like_a_continuing_solver(gale):
    """A solver restarted from a stored solution (gs_solver_set_solution, the
    checkpoint restart) runs the next nonlinear execute bit for bit as the
    solver that produced the solution."""
    from test_gpu_parity import tight
    cfg = _lith3d()
    g = gale.Problem(**cfg.problem_kwargs())
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale, tolerance=1e-11))
    u, p, rep = s.NonLinearExecute(nonLinearTolerance=1e-4, nonLinearMaxIterations=20)
    assert rep.status == 0
    g2 = gale.Problem(**cfg.problem_kwargs())
    s2 = gale.Stokes_SLE_UzawaSolver(g2, tight(gale, tolerance=1e-11))
    s2.set_solution(u, p)
    ua, pa, ra = s.NonLinearExecute(nonLinearTolerance=1e-4, nonLinearMaxIterations=20)
    ub, pb_, rb = s2.NonLinearExecute(nonLinearTolerance=1e-4, nonLinearMaxIterations=20)
    assert ra.status == 0 and rb.status == 0
    assert ra.picard_iterations == rb.picard_iterations
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(pa, pb_)
    with pytest.raises(gale.GaleError):
        s2.set_solution(u[:-1], p)
    s.close()
    s2.close()
    g.close()
    g2.close()

"""StrainWeakening in the oracle, pinned by what the rheology fixes (no GPU).

- With no healing, the increment of a yielding point is the plastic part of
  its strain rate, 2 e_II - tau_y / eta (the rate the viscous strain exceeds
  what the yield stress allows), and zero below yield; with healing and no
  yielding it is -(tau_y / eta) h w.  Checked on linear velocity fields, whose
  strain rate is known exactly at every point.
- A problem whose plastic strain is at or past the final softening strain
  solves exactly as the same problem with the after-softening cohesion and
  friction, and at or below the initial softening strain as the unsoftened one.
- The time step adds dt * increment and clamps at 0; population control
  gives a split the values of a parent in its cell.
"""
import numpy as np
import pytest


def _vm_problem(oracle, res=6, ppc=6, seed=3, C=0.5):
    from gale_amd import configs, layout_random, material, YIELD_VON_MISES
    coord = layout_random(2, (res, res), (0, 0), (1, 1), ppc, seed)
    mats = [material(eta0=1.0, yield_type=YIELD_VON_MISES, cohesion=C)]
    mat = np.zeros(len(coord), np.int32)
    mask, vals = configs.free_slip(2, (res, res))
    return oracle.Problem(2, (res, res), (0, 0), (1, 1), coord, mat, mats, bc_mask=mask, bc_value=vals, gravity=0.0)


def _linear_field(oracle, res, A):
    xs = np.array([oracle.vertex_coord(res, 0.0, 1.0, i) for i in range(res + 1)])
    X, Y = np.meshgrid(xs, xs)
    P = np.stack([X.ravel(), Y.ravel()], 1)   # x fastest
    return (P @ np.asarray(A).T).ravel()


@pytest.mark.parametrize("gamma", [0.1, 0.4, 2.0])
def test_increment_is_plastic_strain_rate(oracle, gamma):
    o = _vm_problem(oracle)
    C, Ca, h = 0.5, 0.2, 0.0
    o.set_strain_weakening(0, healing_rate=h, initial_softening_strain=0.0, final_softening_strain=1.0,
                           cohesion_after=Ca)
    rng = np.random.default_rng(1)
    w = rng.uniform(0.0, 1.5, o.P)
    o.set_plastic_strain(w)
    A = [[0.3 * gamma, gamma], [0.0, -0.3 * gamma]]            # trace-free: deviatoric = full strain rate
    u = _linear_field(oracle, 6, A)
    st, eta, yl = o.viscosity(u, np.zeros(o.E), 1)
    assert st == 0
    S = 0.5 * (np.array(A) + np.array(A).T)
    srII = np.sqrt(0.5 * (S[0, 0] ** 2 + S[1, 1] ** 2) + S[0, 1] ** 2)
    r = np.clip(w, 0.0, 1.0)
    tau_y = C * (1 - r) + Ca * r
    yielding = tau_y < 2.0 * srII
    assert np.array_equal(yl.astype(bool), yielding)
    _, inc = o.plastic_strain()
    assert np.allclose(inc[yielding], 2.0 * srII - tau_y[yielding], rtol=1e-12, atol=1e-14)
    assert np.all(inc[~yielding] == 0.0)
    assert np.allclose(eta[yielding], tau_y[yielding] / (2.0 * srII), rtol=1e-12)   # tau_II = tau_y after yielding
    if gamma == 0.4:
        assert yielding.any() and (~yielding).any()


def test_increment_heals_below_yield(oracle):
    o = _vm_problem(oracle)
    C, Ca, h = 0.5, 0.2, 0.3
    o.set_strain_weakening(0, healing_rate=h, initial_softening_strain=0.1, final_softening_strain=0.9,
                           cohesion_after=Ca)
    w = np.linspace(0.0, 1.2, o.P)
    o.set_plastic_strain(w)
    u = _linear_field(oracle, 6, [[0.0, 0.01], [0.0, 0.0]])    # tau_II = 0.01, far below yield
    st, eta, yl = o.viscosity(u, np.zeros(o.E), 1)
    assert st == 0 and not yl.any()
    r = np.clip((w - 0.1) / 0.8, 0.0, 1.0)
    tau_y = C * (1 - r) + Ca * r
    _, inc = o.plastic_strain()
    assert np.allclose(inc, -tau_y * h * w, rtol=1e-12, atol=1e-15)
    st, _, _ = o.viscosity(u, np.zeros(o.E), 0)                 # no yielding pass (first solve): no increment
    assert np.all(o.plastic_strain()[1] == 0.0)


@pytest.mark.parametrize("state", ["past_final", "before_initial"])
def test_softened_problem_equals_softened_material(oracle, state):
    from gale_amd import configs
    cfg = configs.lithosphere_2d(nx=16, ny=4, ppc=8)
    kw = cfg.oracle_kwargs()
    C, mu = kw["materials"][1]["cohesion"], kw["materials"][1]["friction"]
    Ca, mua = 0.4 * C, 0.5 * mu
    o = oracle.Problem(**kw)
    o.set_strain_weakening(1, healing_rate=0.1, initial_softening_strain=0.5, final_softening_strain=1.5,
                           cohesion_after=Ca, friction_after=mua)
    o.set_plastic_strain(np.full(o.P, 1.5 if state == "past_final" else 0.5))
    kw2 = cfg.oracle_kwargs()
    kw2["materials"] = [dict(m) for m in kw2["materials"]]
    if state == "past_final":
        kw2["materials"][1].update(cohesion=Ca, friction=mua)
    o2 = oracle.Problem(**kw2)
    args = dict(tolerance=1e-8, max_iterations=5000, inner_rtol_override=1e-10, inner_max_iterations=100000)
    st1, u1, p1, r1 = o.picard(oracle.uzawa_params(**args), tolerance=1e-4, max_iterations=20)
    st2, u2, p2, r2 = o2.picard(oracle.uzawa_params(**args), tolerance=1e-4, max_iterations=20)
    assert st1 == st2 == 0
    assert np.array_equal(u1, u2) and np.array_equal(p1, p2)
    w, inc = o.plastic_strain()
    assert np.any(inc != 0.0)


def test_weakening_step_and_population_control(oracle):
    o = _vm_problem(oracle, ppc=10)
    o.set_strain_weakening(0, healing_rate=0.5, final_softening_strain=1.0, cohesion_after=0.1)
    rng = np.random.default_rng(5)
    w0 = rng.uniform(0.0, 1.0, o.P)
    o.set_plastic_strain(w0)
    u = _linear_field(oracle, 6, [[0.09, 0.24], [0.0, -0.09]])   # tau_II = 0.3: weak points yield, others heal
    o.viscosity(u, np.zeros(o.E), 1)
    _, inc = o.plastic_strain()
    assert (inc > 0).any() and (inc < 0).any()
    o.weakening_step(0.7)
    w1, _ = o.plastic_strain()
    assert np.array_equal(w1, np.maximum(w0 + 0.7 * inc, 0.0))
    # population control on an uneven layout: a split carries a parent's values
    owner0 = o.particles()[0]
    vals = np.arange(o.P, dtype=np.float64) * 1e-3
    o.set_plastic_strain(vals)
    before = {e: set(vals[owner0 == e]) for e in range(o.E)}
    P0 = o.P
    added, deleted = o.population_control(upperT=8.0, lowerT=0.5)
    w2, _ = o.plastic_strain()
    owner = o.particles()[0]
    assert o.P == P0 + added - deleted
    for p in range(o.P):
        assert w2[p] in before[owner[p]] or p < P0 - deleted


def test_strain_limited_step(oracle):
    o = _vm_problem(oracle)
    o.set_strain_weakening(0, healing_rate=0.5, final_softening_strain=1.0, cohesion_after=0.1)
    w0 = np.linspace(-0.2, 1.0, o.P)
    o.set_plastic_strain(w0)
    inside = np.arange(o.P) % 3 == 0
    o.weakening_step(0.0, inside=inside, limit=0.25)
    w1, _ = o.plastic_strain()
    assert np.all(w1[inside] == 0.25)
    assert np.array_equal(w1[~inside], np.maximum(w0[~inside], 0.0))

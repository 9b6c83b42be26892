"""Pins the CPU oracle (oracle/gale_oracle.c) against the reference's own
fixtures and against what the mathematics fixes.  CPU only."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN, load_csr, rel

M31 = 2147483647


def gauss_coords(dim, res, lo=(0, 0, 0), hi=(1, 1, 1)):
    """2-point-per-axis Gauss points as global coordinates (x fastest per cell)."""
    g = 1.0 / np.sqrt(3.0)
    E = int(np.prod(res))
    out = []
    for e in range(E):
        ijk = [e % res[0], (e // res[0]) % res[1], e // (res[0] * res[1]) if dim == 3 else 0]
        for q in range(2 ** dim):
            x = [0.0, 0.0, 0.0]
            for d in range(dim):
                a = lo[d] + (ijk[d] / res[d]) * (hi[d] - lo[d])
                b = lo[d] + ((ijk[d] + 1) / res[d]) * (hi[d] - lo[d])
                xi = g if (q >> d) & 1 else -g
                x[d] = a + (1 + xi) * (b - a) / 2
            out.append(x)
    return np.array(out)


def free_slip_mask(dim, res):
    nn = [r + 1 for r in res]
    N = int(np.prod(nn))
    idx = np.indices(nn[::-1]).reshape(dim, -1)[::-1]
    m = np.zeros((N, dim), np.uint8)
    for d in range(dim):
        m[(idx[d] == 0) | (idx[d] == res[d]), d] = 1
    return m.reshape(-1)


def node_coords(dim, res, lo=(0, 0, 0), hi=(1, 1, 1)):
    nn = [r + 1 for r in res]
    idx = np.indices(nn[::-1]).reshape(dim, -1)[::-1].T.astype(float)
    return lo[:dim] + idx / np.array(res) * (np.array(hi[:dim]) - np.array(lo[:dim]))


# ---------------------------------------------------------------------------
def test_park_miller_known_answer(oracle):
    # Park & Miller (1988): x0 = 1 gives x10000 = 1043618065.  The reference
    # XORs the seed with 123459876 (Random.cxx:57-60).
    g = oracle.RngStream(1 ^ 123459876)
    for _ in range(9999):
        g.next()
    assert g.next() == 1043618065.0


def test_random_layout_draws_three_per_particle_in_2d(oracle):
    res, ppc, seed = (3, 2), 5, 13
    coord = oracle.layout_random(2, res, (0.0, 0.0), (1.0, 1.0), ppc, seed)
    g = oracle.RngStream(seed)
    draws = np.array([g.next() for _ in range(3 * coord.shape[0])])
    for p in range(coord.shape[0]):
        e = p // ppc
        i, j = e % res[0], e // res[0]
        x0, x1 = i / 3, (i + 1) / 3
        y0, y1 = j / 2, (j + 1) / 2
        assert coord[p, 0] == (x1 - x0) * draws[3 * p] / M31 + x0
        assert coord[p, 1] == (y1 - y0) * draws[3 * p + 1] / M31 + y0
        assert x0 <= coord[p, 0] <= x1 and y0 <= coord[p, 1] <= y1


@pytest.mark.parametrize("fixture,n", [("k_isoviscous_3x3.npz", 3), ("k_material_viscosity_10x10.npz", 10)])
def test_K_matches_reference_fixture(oracle, fixture, n):
    """Full (BC-free) K of the reference's stiffness-assembly tests: Q1, 2x2
    Gauss, eta = 1 (IsoviscousStiffnessSuite / ViscousSuite, tolerance 1e-4 there)."""
    A = load_csr(fixture).toarray()
    c = gauss_coords(2, (n, n))
    pb = oracle.Problem(2, (n, n), (0, 0), (1, 1), c, np.zeros(len(c), np.int32), [dict(eta0=1.0)])
    assert pb.assemble() == 0
    K = pb.system()[0].toarray()
    assert K.shape == A.shape
    assert rel(K, A) < 1e-13


def test_constitutive_conventions_match_fixture(oracle):
    fx = json.load(open(os.path.join(GOLDEN, "constitutive_2d.json")))
    eta = fx["viscosity"]
    np.testing.assert_array_equal(oracle.isotropic_D(2, eta), np.array(fx["D"]))
    sr = fx["strain_rate"]
    st = oracle.isotropic_stress(2, eta, [sr[0][0], sr[1][1], sr[0][1]])
    exp = fx["stress"]
    np.testing.assert_array_equal(st, [exp[0][0], exp[1][1], exp[0][1]])


def test_K_rigid_modes_and_symmetry_3d(oracle):
    res = (3, 2, 2)
    coord = oracle.layout_random(3, res, (0, 0, 0), (1.5, 1, 1), 7, 5)
    eta0 = 10 ** np.random.default_rng(0).uniform(-2, 2, coord.shape[0])
    pb = oracle.Problem(3, res, (0, 0, 0), (1.5, 1, 1), coord, np.zeros(len(coord), np.int32), [dict(eta0=1.0)],
                        eta0_particle=eta0)
    pb.assemble()
    K = pb.system()[0]
    assert abs(K - K.T).max() < 1e-12 * abs(K).max()
    X = node_coords(3, res, hi=(1.5, 1, 1))
    modes = [np.tile(np.eye(3)[d], (len(X), 1)) for d in range(3)]
    modes += [np.stack([-X[:, 1], X[:, 0], 0 * X[:, 0]], 1), np.stack([-X[:, 2], 0 * X[:, 0], X[:, 0]], 1),
              np.stack([0 * X[:, 0], -X[:, 2], X[:, 1]], 1)]
    for m in modes:
        y = K @ m.reshape(-1)
        assert np.linalg.norm(y) < 1e-12 * abs(K).max() * np.linalg.norm(m)
    # a non-rigid field is not annihilated (catches a zero operator)
    y = K @ (X[:, [0, 0, 0]] ** 2).reshape(-1)
    assert np.linalg.norm(y) > 1e-3


@pytest.mark.parametrize("dim", [2, 3])
def test_G_divergence_theorem(oracle, dim):
    """(G^T u)_e = -int_e div u for a linear u = A x (exact on Q1)."""
    res = (3, 4) if dim == 2 else (2, 3, 2)
    hi = (1.2, 0.8, 1.0)
    coord = oracle.layout_random(dim, res, (0, 0, 0), hi, 3, 7)
    pb = oracle.Problem(dim, res, (0, 0, 0), hi, coord, np.zeros(len(coord), np.int32), [dict(eta0=1.0)])
    A = np.random.default_rng(1).normal(size=(dim, dim))
    X = node_coords(dim, res, hi=hi)
    u = (X @ A.T).reshape(-1)
    q = pb.apply_Gt(u)
    vol = np.prod([hi[d] / res[d] for d in range(dim)])
    np.testing.assert_allclose(q, -np.trace(A) * vol, rtol=1e-12)
    # and G = (G^T)^T
    p = np.random.default_rng(2).normal(size=pb.E)
    x = np.random.default_rng(3).normal(size=pb.N * dim)
    assert abs(pb.apply_G(p) @ x - p @ pb.apply_Gt(x)) < 1e-12 * np.linalg.norm(p) * np.linalg.norm(x)


def _sinker(oracle, dim, res, ppc=12, seed=13):
    hi = (1.0, 1.0, 1.0)
    coord = oracle.layout_random(dim, res, (0, 0, 0), hi, ppc, seed)
    c = np.array([0.5, 0.6, 0.5][:dim])
    mat = (((coord[:, :dim] - c) ** 2).sum(1) <= 0.2 ** 2).astype(np.int32)
    mats = [dict(eta0=1.0, density=1.0), dict(eta0=100.0, density=1.3)]
    return oracle.Problem(dim, res, (0, 0, 0), hi, coord, mat, mats, bc_mask=free_slip_mask(dim, res))


@pytest.mark.parametrize("dim,res", [(2, (8, 8)), (3, (3, 3, 3))])
def test_uzawa_matches_direct_saddle_solve(oracle, dim, res):
    pb = _sinker(oracle, dim, res)
    pb.assemble()
    K, G, f, h, Q, eq = pb.system()
    A = sp.bmat([[K, G], [G.T, None]]).toarray()
    sol = np.linalg.lstsq(A, np.concatenate([f, h]), rcond=None)[0]
    nu = K.shape[0]
    ud, pd = sol[:nu], sol[nu:]
    for inner in (0, 1):
        st, u, p, rep = pb.uzawa(tolerance=1e-11, inner=inner, inner_rtol_override=1e-14, max_iterations=5000)
        assert st == 0
        uf = u[eq >= 0]
        assert rel(uf, ud) < 1e-8
        assert rel(p - p.mean(), pd - pd.mean()) < 1e-8
        assert rep.div_u < 1e-8 and rep.momentum_residual < 1e-9


def test_hydrostatic_closed_form(oracle):
    """Uniform density in a free-slip box: u = 0 and p = rho g (H - y) + c
    exactly on Q1-P0 with exact (Gauss) integration."""
    res = (6, 5)
    c = gauss_coords(2, res)
    pb = oracle.Problem(2, res, (0, 0), (1, 1), c, np.zeros(len(c), np.int32), [dict(eta0=3.0, density=2.0)],
                        bc_mask=free_slip_mask(2, res), gravity=1.5)
    st, u, p, rep = pb.uzawa(tolerance=1e-12, inner_rtol_override=1e-15, max_iterations=2000)
    assert st == 0
    yc = (np.arange(pb.E) // res[0] + 0.5) / res[1]
    exact = 2.0 * 1.5 * (1.0 - yc)
    assert np.abs(u).max() < 1e-10
    np.testing.assert_allclose(p - p.mean(), exact - exact.mean(), atol=1e-9)


def _shear_problem(oracle, mats, gamma, p_elem=None):
    res = (4, 4)
    coord = oracle.layout_random(2, res, (0, 0), (1, 1), 6, 3)
    pb = oracle.Problem(2, res, (0, 0), (1, 1), coord, np.zeros(len(coord), np.int32), mats)
    X = node_coords(2, res)
    u = np.stack([gamma * X[:, 1], 0 * X[:, 0]], 1).reshape(-1)
    return pb, u, (np.full(pb.E, p_elem) if p_elem is not None else None)


def test_von_mises_simple_shear_closed_form(oracle):
    # simple shear u = (gamma y, 0): strain-rate invariant gamma/2, stress
    # invariant eta0*gamma; above the yield stress C the viscosity is C/gamma.
    gamma, eta0, C = 2.0, 5.0, 3.0
    pb, u, _ = _shear_problem(oracle, [dict(eta0=eta0, yield_type=2, cohesion=C)], gamma)
    st, eta, yl = pb.viscosity(u, None, yield_on=1)
    np.testing.assert_allclose(eta, C / gamma, rtol=1e-14)
    assert yl.all()
    st, eta, yl = pb.viscosity(u, None, yield_on=0)  # first solve: no yielding (YieldRheology.cxx:203-204)
    np.testing.assert_array_equal(eta, eta0)
    pb2, u2, _ = _shear_problem(oracle, [dict(eta0=eta0, yield_type=2, cohesion=20.0)], gamma)
    assert (pb2.viscosity(u2, None, 1)[1] == eta0).all()  # below yield: unchanged


def test_drucker_prager_2d_closed_form(oracle):
    # 2D: tau_y = (mu p + C) cos(phi) with cos(phi) = 1/sqrt(1+mu^2)
    gamma, eta0, C, mu, pres = 4.0, 10.0, 0.5, np.tan(np.radians(30)), 2.0
    mats = [dict(eta0=eta0, yield_type=1, cohesion=C, friction=mu, min_viscosity=1e-3)]
    pb, u, p = _shear_problem(oracle, mats, gamma, pres)
    st, eta, yl = pb.viscosity(u, p, 1)
    tau_y = (mu * pres + C) / np.sqrt(1 + mu * mu)
    np.testing.assert_allclose(eta, tau_y / gamma, rtol=1e-13)
    # minimum viscosity floor
    mats[0]["min_viscosity"] = 1.0
    pb, u, p = _shear_problem(oracle, mats, gamma, pres)
    np.testing.assert_allclose(pb.viscosity(u, p, 1)[1], 1.0)
    # tension below the minimum yield stress: clamp at C cos(phi)
    mats[0]["min_viscosity"] = 0.0
    pb, u, p = _shear_problem(oracle, mats, gamma, -100.0)
    np.testing.assert_allclose(pb.viscosity(u, p, 1)[1], C / np.sqrt(1 + mu * mu) / gamma, rtol=1e-13)


def test_drucker_prager_3d_closed_form(oracle):
    # 3D simple shear u = (gamma y, 0, 0): strain-rate invariant gamma/2,
    # stress invariant eta gamma.  DruckerPrager.cxx:376-405 in 3D:
    # f = 2 cos(phi) / (sqrt(3) (3 - sin(phi))), tau_y = mu f p + 3 C f,
    # clamped below at minimumYieldStress f (minimumYieldStress = C when 0);
    # the yielded viscosity is tau_y / (2 eps_II) = tau_y / gamma.
    gamma, eta0, C, mu, pres = 4.0, 10.0, 0.5, np.tan(np.radians(30)), 2.0
    cphi = 1.0 / np.sqrt(1.0 + mu * mu)
    sphi = mu * cphi
    f = 2.0 * cphi / (np.sqrt(3.0) * (3.0 - sphi))
    res = (3, 3, 2)
    coord = oracle.layout_random(3, res, (0, 0, 0), (1, 1, 1), 5, 7)

    def visc(mats, p_elem):
        pb = oracle.Problem(3, res, (0, 0, 0), (1, 1, 1), coord, np.zeros(len(coord), np.int32), mats)
        X = node_coords(3, res)
        u = np.stack([gamma * X[:, 1], 0 * X[:, 0], 0 * X[:, 0]], 1).reshape(-1)
        return pb.viscosity(u, np.full(pb.E, p_elem), 1)

    mats = [dict(eta0=eta0, yield_type=1, cohesion=C, friction=mu, min_viscosity=1e-3)]
    st, eta, yl = visc(mats, pres)
    tau_y = mu * f * pres + 3.0 * C * f
    assert tau_y < eta0 * gamma and yl.all()
    np.testing.assert_allclose(eta, tau_y / gamma, rtol=1e-13)
    # the 2D criterion would give a different value: the 3D branch is the one taken
    assert abs(tau_y - (mu * pres + C) * cphi) > 1e-3
    # tension below the minimum yield stress: clamp at C f (no factor 3)
    st, eta, yl = visc(mats, -100.0)
    np.testing.assert_allclose(eta, C * f / gamma, rtol=1e-13)
    # an explicit minimumYieldStress replaces C in the clamp
    mats2 = [dict(mats[0], min_yield_stress=0.8)]
    st, eta, yl = visc(mats2, -100.0)
    np.testing.assert_allclose(eta, 0.8 * f / gamma, rtol=1e-13)
    # minimum viscosity floor
    mats3 = [dict(mats[0], min_viscosity=1.0)]
    np.testing.assert_allclose(visc(mats3, pres)[1], 1.0)
    # below yield: unchanged
    mats4 = [dict(mats[0], cohesion=50.0)]
    st, eta, yl = visc(mats4, pres)
    np.testing.assert_array_equal(eta, eta0)
    assert not yl.any()


def _q_closed_form(dim, h, eta):
    """UzawaPreconditionerTerm.cxx:195-215 on one Q1-P0 element with exact
    integration: Q_e = sum over the element's velocity dofs v of
    G_e[v]^2 / K_e[v][v] (the element's own K diagonal).  On a box of sides
    h, trilinear N_a: |int dN_a/dx_d| = prod_{c != d} h_c / 2 and
    int (dN_a/dx_c)^2 = prod_{c' != c} (h_c' / 3) / h_c, with the isotropic
    D = eta diag(2, .., 2, 1, .., 1): K_dd = eta sum_c (1 + [c == d]) int (dN_a/dx_c)^2."""
    h = np.asarray(h, float)
    q = 0.0
    for d in range(dim):
        area = np.prod([h[c] for c in range(dim) if c != d])
        g2 = (area / 2 ** (dim - 1)) ** 2
        kdd = 0.0
        for c in range(dim):
            ic = np.prod([h[k] / 3.0 for k in range(dim) if k != c]) / h[c]
            kdd += (2.0 if c == d else 1.0) * ic
        q += 2 ** dim * g2 / (eta * kdd)
    return q


@pytest.mark.parametrize("dim", [2, 3])
def test_uzawa_Q_closed_form(oracle, dim):
    """The oracle's Uzawa preconditioner equals the element-level closed form
    on every element -- also on boundary elements with Dirichlet dofs, where
    an assembled (global) diagonal of K or dropping the Dirichlet dofs would
    give other values -- with an element-wise viscosity and unequal sides."""
    res = (4, 3, 2)[:dim]
    hi = (1.0, 0.6, 0.35)[:dim]
    c = gauss_coords(dim, res, (0, 0, 0), hi + (1.0,) * (3 - dim))
    E = int(np.prod(res))
    eta_e = np.array([1.0 + 0.5 * e for e in range(E)])
    mat = np.repeat(np.arange(E, dtype=np.int32), 2 ** dim)
    mats = [dict(eta0=float(v), density=1.0) for v in eta_e]
    pb = oracle.Problem(dim, res, (0,) * dim, hi, c, mat, mats, bc_mask=free_slip_mask(dim, res))
    pb.assemble()
    q = pb.uzawa_Q()
    h = [hi[d] / res[d] for d in range(dim)]
    exact = np.array([_q_closed_form(dim, h, v) for v in eta_e])
    np.testing.assert_allclose(q, exact, rtol=1e-12)
    # the square / cube special cases: 2 h^2 / eta (2D), 27 h^3 / (8 eta) (3D)
    assert abs(_q_closed_form(2, (0.5, 0.5), 2.0) - 2 * 0.25 / 2.0) < 1e-15
    assert abs(_q_closed_form(3, (0.5,) * 3, 2.0) - 27 * 0.125 / 16.0) < 1e-15


def test_ownership_rule_on_face_lattice(oracle):
    """Mesh_RegularAlgorithms rule: floor((x-min)/sep); a point whose
    fractional part (in element units) is within 1e-12 of a face goes to the
    lower element; within 1e-12 of max goes to the last element; outside -> none."""
    res, lo, hi = (8, 4), (0.0, 0.0), (1.0, 0.5)   # sep = 0.125 exactly
    search = lambda x: oracle.search_element(2, res, lo, hi, x)
    for i in range(9):
        for j in range(5):
            x, y = i * 0.125, j * 0.125
            ei = min(max(i - 1, 0), 7) if i > 0 else 0
            ej = min(max(j - 1, 0), 3) if j > 0 else 0
            assert search((x, y)) == ei + 8 * ej
    # 1e-13 element units above an interior face still goes to the lower element
    assert search((3 * 0.125 + 1e-14, 0.01)) == 2
    # 1e-11 element units above goes to the upper one
    assert search((3 * 0.125 + 0.125 * 1e-11, 0.01)) == 3
    # just below a face
    assert search((np.nextafter(3 * 0.125, 0.0), 0.01)) == 2
    # max face tolerance is in coordinate units
    assert search((1.0 + 5e-13, 0.01)) == 7
    assert search((1.0 + 2e-12, 0.01)) == -1
    assert search((-1e-300, 0.01)) == -1
    assert search((0.0, 0.5)) == 8 * 3
    # interior points: plain floor
    rng = np.random.default_rng(4)
    for _ in range(200):
        x, y = rng.uniform(0, 1), rng.uniform(0, 0.5)
        fi, fj = int(x / 0.125), int(y / 0.125)
        if min(abs(x / 0.125 - round(x / 0.125)), abs(y / 0.125 - round(y / 0.125))) < 1e-9:
            continue
        assert search((x, y)) == fi + 8 * fj


def test_rk2_exact_on_linear_field(oracle):
    """Midpoint RK2 on a linear velocity (represented exactly by Q1):
    x1 = (I + dt A + dt^2 A^2 / 2) x0 (TimeIntegrand_SecondOrder)."""
    res = (8, 8)
    coord = oracle.layout_random(2, res, (0, 0), (1, 1), 4, 9)
    pb = oracle.Problem(2, res, (0, 0), (1, 1), coord, np.zeros(len(coord), np.int32), [dict(eta0=1.0)])
    A = np.array([[0.0, -1.0], [1.0, 0.0]])  # rigid rotation about (0.5, 0.5)
    X = node_coords(2, res)
    u = ((X - 0.5) @ A.T).reshape(-1)
    dt = 0.05
    st, out = pb.advect_rk2(u, dt)
    assert st == 0
    x0 = coord[:, :2] - 0.5
    M = np.eye(2) + dt * A + 0.5 * dt * dt * A @ A
    xm = x0 + 0.5 * dt * x0 @ A.T + 0.5
    outside = ((xm < 0) | (xm > 1)).any(1)   # midpoint left the box: first-order fallback
    exp = np.where(outside[:, None], x0 + dt * x0 @ A.T, x0 @ M.T)
    assert 0 < outside.sum() < len(outside)
    np.testing.assert_allclose(out[:, :2] - 0.5, exp, atol=1e-14)
    # dt = factor * 0.5 * minSep / max|v|  (UpdateDt.cxx:53-84)
    vmax = np.sqrt(((u.reshape(-1, 2)) ** 2).sum(1)).max()
    assert pb.compute_dt(u, 1.0) == pytest.approx(0.5 * 0.125 / vmax, rel=1e-15)


def test_picard_linear_problem_converges_on_second_solve(oracle):
    pb = _sinker(oracle, 2, (6, 6))
    up = oracle.uzawa_params(tolerance=1e-11, inner_rtol_override=1e-14, max_iterations=5000)
    st, u, p, rep = pb.picard(up, tolerance=1e-6, max_iterations=10)
    assert st == 0 and rep.picard_iterations == 2 and rep.picard_residual < 1e-9


def test_uzawa_firewalls(oracle):
    res = (4, 4)
    coord = oracle.layout_random(2, res, (0, 0), (1, 1), 4, 1)
    pb = oracle.Problem(2, res, (0, 0), (1, 1), coord, np.zeros(len(coord), np.int32),
                        [dict(eta0=1.0, density=0.0)], bc_mask=free_slip_mask(2, res))
    st, u, p, rep = pb.uzawa(tolerance=1e-6)
    assert st == oracle.E_ZERO_FORCE
    pb = _sinker(oracle, 2, (6, 6))
    st, u, p, rep = pb.uzawa(tolerance=1e-14, max_iterations=1)
    assert st == oracle.E_MAXIT


# ---------------------------------------------------------------------------
# Output fields of the material points (StoreStress / StoreViscosity / the
# strain-rate field): a linear velocity field has the constant strain rate
# sym(A) everywhere, whose invariant and isotropic stress are closed forms.
@pytest.mark.parametrize("dim", [2, 3])
def test_particle_fields_linear_velocity_closed_form(oracle, dim):
    from gale_amd import configs
    cfg = configs.sinker_2d(res=6, ppc=6, seed=4) if dim == 2 else configs.rayleigh_taylor_3d(res=4, ppc=4, seed=4)
    o = oracle.Problem(**cfg.oracle_kwargs())
    o.viscosity()
    rng = np.random.default_rng(dim)
    A = rng.normal(size=(dim, dim))
    b = rng.normal(size=dim)
    res, lo, hi = list(cfg.res), list(cfg.lo), list(cfg.hi)
    axes = [np.array([oracle.vertex_coord(res[d], lo[d], hi[d], i) for i in range(res[d] + 1)]) for d in range(dim)]
    grids = np.meshgrid(*axes, indexing="ij")
    X = np.stack([g.transpose().ravel() for g in grids], axis=1)  # x fastest
    u = (X @ A.T + b).ravel()
    f = o.particle_fields(u)
    S = 0.5 * (A + A.T)
    if dim == 2:
        want = np.array([S[0, 0], S[1, 1], S[0, 1]])
    else:
        want = np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])
    assert np.allclose(f["strain_rate"], want, rtol=1e-12, atol=1e-12)
    inv = np.sqrt(0.5 * np.sum(S * S))
    assert np.allclose(f["strain_rate_inv"], inv, rtol=1e-12)
    eta = f["viscosity"][:, None]
    assert np.allclose(f["stress"], 2.0 * eta * want, rtol=1e-12, atol=1e-12)

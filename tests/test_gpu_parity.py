"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Integer/index work is compared bit for bit; FP64
operators to 1e-12 relative; solves at tight tolerances to 1e-8 relative L2
(pressure modulo a constant on closed boxes)."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def full_from_reduced(red, eq):
    out = np.zeros(eq.size)
    out[eq >= 0] = red
    return out


def pair(gale, oracle, cfg):
    g = gale.Problem(**cfg.problem_kwargs())
    o = oracle.Problem(**cfg.oracle_kwargs())
    return g, o


def small_sinker(dim, res, ppc=8, seed=13):
    from gale_amd import configs
    if dim == 2:
        return configs.sinker_2d(res=res, ppc=ppc, seed=seed)
    cfg = configs.rayleigh_taylor_3d(res=res, ppc=ppc, seed=seed)
    return cfg


def tight(gale, **kw):
    args = dict(tolerance=1e-10, inner_rtol_override=1e-13, maxIterations=20000, inner_max_iterations=200000)
    args.update(kw)
    return gale.uzawa_params(**args)


def otight(oracle, **kw):
    args = dict(tolerance=1e-10, inner_rtol_override=1e-13, max_iterations=20000, inner_max_iterations=200000)
    args.update(kw)
    return oracle.uzawa_params(**args)


# ---------------------------------------------------------------------------
def test_relocation_bit_exact_on_ties(gale, oracle):
    """Owner indices from the GPU search + sort equal the reference rule on
    random points, points on every face/edge/corner and +-1e-13 / +-1 ulp off."""
    res, lo, hi = (8, 4), (0.0, 0.0), (1.0, 0.3)
    rng = np.random.default_rng(7)
    pts = [np.c_[rng.uniform(0, 1, 400), rng.uniform(0, 0.3, 400), np.zeros(400)]]
    xs = [lo[0] + (i / 8) * (hi[0] - lo[0]) for i in range(9)]
    ys = [lo[1] + (j / 4) * (hi[1] - lo[1]) for j in range(5)]
    lat = []
    for x in xs:
        for y in ys:
            for dx in (0.0, 1e-13, -1e-13, 1e-17):
                for ex in (0, 1, -1):
                    xx = np.nextafter(x, 2.0) if ex == 1 else (np.nextafter(x, -2.0) if ex == -1 else x)
                    lat.append((min(max(xx + dx, 0.0), 1.0), min(max(y, 0.0), 0.3), 0.0))
    pts.append(np.array(lat))
    coord = np.concatenate(pts)
    mat = np.zeros(len(coord), np.int32)
    g = gale.Problem(2, res, lo, hi, coord, mat, [gale.material()])
    _, owner, xi, w = g.particles()
    expected = np.array([oracle.search_element(2, res, lo, hi, c[:2]) for c in coord])
    np.testing.assert_array_equal(owner, expected)
    o = oracle.Problem(2, res, lo, hi, coord, mat, [dict(eta0=1.0)])
    oo, oxi, ow, _ = o.particles()
    np.testing.assert_array_equal(owner, oo)
    np.testing.assert_array_equal(w, ow)
    assert np.abs(xi - oxi).max() < 1e-13
    # a second relocation is deterministic
    st, owner2 = g.relocate()
    np.testing.assert_array_equal(owner2, owner)


@pytest.mark.parametrize("dim,res", [(2, 12), (3, 4)])
def test_setup_arrays_match_oracle(gale, oracle, dim, res):
    cfg = small_sinker(dim, res)
    g, o = pair(gale, oracle, cfg)
    s = gale.Stokes_SLE_UzawaSolver(g, tolerance=1e-6)
    s.setup()
    a = s.setup_arrays()
    o.assemble()
    K, G, f, h, Q, eq = o.system()
    assert rel(a["diagK"], o.diag_K()) < 1e-13
    assert rel(a["Q"], Q) < 1e-12
    assert rel(a["f"], full_from_reduced(f, eq)) < 1e-12
    assert rel(a["h"], h) < 1e-12 or np.abs(h).max() == 0
    st, eta, _ = o.viscosity()
    np.testing.assert_array_equal(a["eta"], eta)


@pytest.mark.parametrize("dim,res", [(2, 10), (3, 4)])
def test_operator_applies_match_oracle(gale, oracle, dim, res):
    import torch
    from gale_amd import configs
    cfg = configs.random_viscosity_3d(res=res, ppc=6) if dim == 3 else configs.sinker_2d(res=res, ppc=9)
    g, o = pair(gale, oracle, cfg)
    s = gale.Stokes_SLE_UzawaSolver(g)
    s.setup()
    o.assemble()
    rng = np.random.default_rng(11)
    x = rng.normal(size=g.N * dim)
    p = rng.normal(size=g.E)
    dev = torch.device("cuda:0")
    xd = torch.tensor(x, device=dev)
    pd = torch.tensor(p, device=dev)
    y = torch.zeros_like(xd)
    y2 = torch.zeros_like(xd)
    gy = torch.zeros_like(xd)
    q = torch.zeros_like(pd)
    torch.cuda.synchronize()
    s.apply_K(xd, y)
    s.apply_K(xd, y2)
    s.apply_G(pd, gy)
    s.apply_Gt(xd, q)
    s.sync()
    yk = y.cpu().numpy()
    assert rel(yk, o.apply_K(x)) < 1e-13
    np.testing.assert_array_equal(yk, y2.cpu().numpy())  # deterministic scatter
    assert rel(gy.cpu().numpy(), o.apply_G(p)) < 1e-13
    assert rel(q.cpu().numpy(), o.apply_Gt(x)) < 1e-13
    # moment form of K (what the solver applies): the same operator up to rounding
    ym = torch.zeros_like(xd)
    torch.cuda.synchronize()
    s.apply_K_moments(xd, ym)
    s.sync()
    assert rel(ym.cpu().numpy(), yk) < 1e-12
    # FP32 apply (the mixed-precision preconditioner)
    x32 = xd.float()
    y32 = torch.zeros_like(x32)
    torch.cuda.synchronize()
    s.apply_K_fp32(x32, y32)
    s.sync()
    assert rel(y32.double().cpu().numpy(), yk) < 2e-5


@pytest.mark.parametrize("dim,res,inner,prec", [(2, 16, 1, 0), (3, 5, 1, 0), (2, 16, 0, 0), (3, 6, 0, 0),
                                                (2, 16, 1, 1), (3, 8, 1, 1)])
def test_uzawa_parity_tight(gale, oracle, dim, res, inner, prec):
    """Jacobi-CG, multigrid-FCG and the FP32-V-cycle mixed form all reach the
    oracle's discrete solution at tight tolerances."""
    cfg = small_sinker(dim, res)
    g, o = pair(gale, oracle, cfg)
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale, inner_preconditioner=inner, precision=prec))
    u, p, rep = s.Solve()
    st, ou, op, orep = o.uzawa(otight(oracle))
    assert st == 0 and rep.status == 0
    assert rel(u, ou) < 1e-8
    assert rel(p - p.mean(), op - op.mean()) < 1e-8
    assert rep.div_u < 1e-7 and rep.momentum_residual < 1e-9
    # same method, same outer iteration count (inner solves converged tightly)
    assert abs(rep.uzawa_iterations - orep.uzawa_iterations) <= max(2, orep.uzawa_iterations // 20)


def test_sinker_reference_tolerance(gale, oracle):
    """The 2D sinker at full size (32^2, 20 ppc) with the reference settings:
    Uzawa 1e-6 relative residual, inner rtol 0.1 x tolerance."""
    from gale_amd import configs
    cfg = configs.sinker_2d()
    g, o = pair(gale, oracle, cfg)
    s = gale.Stokes_SLE_UzawaSolver(g, tolerance=1e-6)
    u, p, rep = s.Solve()
    assert rep.status == 0 and rep.rel_residual <= 1e-6
    st, ou, op, orep = o.uzawa(otight(oracle))
    assert st == 0
    assert rel(u, ou) < 1e-4
    assert rel(p - p.mean(), op - op.mean()) < 1e-4
    # one RK2 step with dt = 0.5 minSep / max|v| (dtFactor 1)
    dt = s.compute_dt(1.0)
    assert dt == pytest.approx(o.compute_dt(u, 1.0), rel=1e-14)
    before = g.particles()[0]
    st = s.advect_rk2(dt)
    assert st == 0
    after, owner, _, _ = g.particles()
    ost, oafter = o.advect_rk2(u, dt)
    assert ost == 0
    assert np.abs(after - oafter).max() < 1e-14
    assert np.abs(after - before).max() > 0
    # ownership of identical positions is bit-exact
    o.set_coords(after)
    np.testing.assert_array_equal(owner, o.particles()[0])


def test_hydrostatic_on_gpu(gale, oracle):
    res = (6, 5)
    coord = gale.layout_gauss(2, res, (0, 0), (1, 1))
    mask, vals = gale.free_slip(2, res)
    g = gale.Problem(2, res, (0, 0), (1, 1), coord, np.zeros(len(coord), np.int32),
                     [gale.material(eta0=3.0, density=2.0)], bc_mask=mask, bc_value=vals, gravity=1.5)
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale, tolerance=1e-12, inner_rtol_override=1e-15))
    u, p, rep = s.Solve()
    yc = (np.arange(g.E) // res[0] + 0.5) / res[1]
    exact = 2.0 * 1.5 * (1.0 - yc)
    assert np.abs(u).max() < 1e-10
    np.testing.assert_allclose(p - p.mean(), exact - exact.mean(), atol=1e-9)


def test_picard_drucker_prager_parity(gale, oracle):
    from gale_amd import configs
    cfg = configs.lithosphere_2d(nx=32, ny=8, ppc=10)
    g, o = pair(gale, oracle, cfg)
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale, tolerance=1e-11))
    u, p, rep = s.NonLinearExecute(nonLinearTolerance=1e-4, nonLinearMaxIterations=20)
    st, ou, op, orep = o.picard(otight(oracle, tolerance=1e-11), tolerance=1e-4, max_iterations=20)
    assert st == 0 and rep.status == 0
    assert rep.picard_iterations == orep.picard_iterations
    assert rel(u, ou) < 1e-8
    assert rel(p, op) < 1e-8
    eta = s.setup_arrays()["eta"]
    assert eta.min() < 1e2 * (1 - 1e-6)  # the lid yields somewhere


def test_get_rhs_and_form_residual(gale, oracle):
    cfg = small_sinker(2, 10)
    g = gale.Problem(**cfg.problem_kwargs())
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale))
    u, p, rep = s.Solve()
    fhat = s.GetRhs()
    r = s.FormResidual(p)
    assert np.linalg.norm(r) < 1e-8 * np.linalg.norm(fhat)
    r0 = s.FormResidual(np.zeros_like(p))
    assert rel(r0, fhat) < 1e-12


def test_firewalls(gale):
    res = (4, 4)
    coord = gale.layout_random(2, res, (0, 0), (1, 1), 4, 1)
    mask, vals = gale.free_slip(2, res)
    g = gale.Problem(2, res, (0, 0), (1, 1), coord, np.zeros(len(coord), np.int32),
                     [gale.material(density=0.0)], bc_mask=mask, bc_value=vals)
    s = gale.Stokes_SLE_UzawaSolver(g, tolerance=1e-6)
    with pytest.raises(gale.GaleError) as e:
        s.Solve()
    assert e.value.status == gale.E_ZERO_FORCE
    cfg = small_sinker(2, 6)
    g = gale.Problem(**cfg.problem_kwargs())
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale, tolerance=1e-14, maxIterations=1))
    with pytest.raises(gale.GaleError) as e:
        s.Solve()
    assert e.value.status == gale.E_MAXIT

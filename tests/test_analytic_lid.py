"""Isoviscous Stokes flow under a sinusoidal lid against its closed form.

The reference checks its solver on this flow (StgFEM_LidDrivenIsoviscousAnalytic,
StgFEM/SysTest/AnalyticPlugins/LidDrivenIsoviscousAnalytic/).  Here the closed
form is derived from the mathematics, not taken from that plugin: with the
stream function psi = sin(k x) g(y), k = n pi, the biharmonic equation gives
g = (A + C y) e^{ky} + (B + D y) e^{-ky}; u = psi_y, v = -psi_x; and the
four conditions g(0) = g'(0) = 0 (no slip at the base), g(1) = 0, g'(1) = 1
(u = sin(k x), v = 0 on the lid) fix A..D.  The sides x = 0, 1 are free
slip (u = 0, no shear) as the solution requires.  Stokes with eta = 1 gives
p = -2 k cos(k x) (C e^{ky} + D e^{-ky}) up to a constant.  Q1-P0 velocities
must converge at second order and the pressure at first order at least."""
import numpy as np
import pytest


def closed_form(n):
    k = n * np.pi
    # rows: g(0), g'(0), g(1), g'(1) in the unknowns A, B, C, D
    e, ei = np.exp(k), np.exp(-k)
    M = np.array([[1.0, 1.0, 0.0, 0.0],
                  [k, -k, 1.0, 1.0],
                  [e, ei, e, ei],
                  [k * e, -k * ei, (1 + k) * e, (1 - k) * ei]])
    A, B, C, D = np.linalg.solve(M, [0.0, 0.0, 0.0, 1.0])

    def vel(x, y):
        gp = (k * A + C + k * C * y) * np.exp(k * y) + (D - k * B - k * D * y) * np.exp(-k * y)
        g = (A + C * y) * np.exp(k * y) + (B + D * y) * np.exp(-k * y)
        return np.sin(k * x) * gp, -k * np.cos(k * x) * g

    def pres(x, y):
        return -2.0 * k * np.cos(k * x) * (C * np.exp(k * y) + D * np.exp(-k * y))

    return vel, pres


def setup(G, res, n=1):
    vel, pres = closed_form(n)
    r = (res, res)
    coord = G.layout_gauss(2, r, (0, 0), (1, 1))
    mask, vals = G.wall_bc(2, r, [("bottom", 0, 0.0), ("bottom", 1, 0.0), ("left", 0, 0.0), ("right", 0, 0.0),
                                  ("top", 1, 0.0)])
    xs = np.linspace(0.0, 1.0, res + 1)
    for i in range(res + 1):   # lid: u = sin(k x), v = 0
        node = res * (res + 1) + i
        mask[2 * node] = 1
        vals[2 * node] = vel(xs[i], 1.0)[0]
    X, Y = np.meshgrid(xs, xs)
    ue = np.stack(vel(X.ravel(), Y.ravel()), 1)
    xc = 0.5 * (xs[1:] + xs[:-1])
    XC, YC = np.meshgrid(xc, xc)
    pe = pres(XC.ravel(), YC.ravel())
    return coord, mask, vals, ue, pe


def errors(u, p, ue, pe):
    eu = np.linalg.norm(u.reshape(-1, 2) - ue) / np.linalg.norm(ue)
    pm, pem = p - p.mean(), pe - pe.mean()
    return eu, np.linalg.norm(pm - pem) / np.linalg.norm(pem)


def test_oracle_converges_to_closed_form(oracle):
    import gale_amd as G
    errs = []
    for res in (8, 16, 32):
        coord, mask, vals, ue, pe = setup(G, res)
        o = oracle.Problem(2, (res, res), (0, 0), (1, 1), coord, np.zeros(len(coord), np.int32),
                           [dict(eta0=1.0, density=0.0)], bc_mask=mask, bc_value=vals, gravity=0.0)
        st, u, p, _ = o.uzawa(oracle.uzawa_params(tolerance=1e-10, max_iterations=5000, inner_rtol_override=1e-12,
                                                  inner_max_iterations=100000))
        assert st == 0
        errs.append(errors(u, p, ue, pe))
    print(errs)
    for (eu0, ep0), (eu1, ep1) in zip(errs, errs[1:]):
        assert eu0 / eu1 > 3.5 and ep0 / ep1 > 1.8
    assert errs[-1][0] < 0.015 and errs[-1][1] < 0.15     # the reference plugin's tolerances


@pytest.mark.gpu
def test_gpu_converges_to_closed_form(gale):
    errs = []
    for res in (16, 32, 64, 128):
        coord, mask, vals, ue, pe = setup(gale, res)
        pb = gale.Problem(dim=2, res=(res, res), lo=(0, 0), hi=(1, 1), coord=coord,
                          material_index=np.zeros(len(coord), np.int32), materials=[gale.material(eta0=1.0)],
                          bc_mask=mask, bc_value=vals, gravity=0.0)
        s = gale.Stokes_SLE_UzawaSolver(pb, gale.uzawa_params(tolerance=1e-10, maxIterations=10000,
                                                             inner_rtol_override=1e-12))
        u, p, rep = s.Solve()
        assert rep.status == 0
        errs.append(errors(u, p, ue, pe))
        print(f"res {res}: velocity {errs[-1][0]:.3e} pressure {errs[-1][1]:.3e}", flush=True)
        s.close()
        pb.close()
    for (eu0, ep0), (eu1, ep1) in zip(errs, errs[1:]):
        assert eu0 / eu1 > 3.5 and ep0 / ep1 > 1.8
    assert errs[1][0] < 0.015 and errs[1][1] < 0.15


def setup3(G, res, n=1):
    """The same flow in a unit cube, independent of z: w = 0, free slip on
    the z walls (no normal velocity, no shear), no slip at the base y = 0,
    the lid y = 1 moving with u = sin(k x)."""
    vel, pres = closed_form(n)
    r = (res, res, res)
    coord = G.layout_gauss(3, r, (0, 0, 0), (1, 1, 1))
    walls = [("bottom", 0, 0.0), ("bottom", 1, 0.0), ("bottom", 2, 0.0), ("left", 0, 0.0), ("right", 0, 0.0),
             ("top", 1, 0.0), ("top", 2, 0.0), ("back", 2, 0.0), ("front", 2, 0.0)]
    mask, vals = G.wall_bc(3, r, walls)
    xs = np.linspace(0.0, 1.0, res + 1)
    nn = res + 1
    for kz in range(nn):
        for i in range(nn):
            node = i + nn * (res + nn * kz)
            mask[3 * node] = 1
            vals[3 * node] = vel(xs[i], 1.0)[0]
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    Xn, Yn = X.transpose(2, 1, 0).ravel(), Y.transpose(2, 1, 0).ravel()   # x fastest
    ux, uy = vel(Xn, Yn)
    ue = np.stack([ux, uy, np.zeros_like(ux)], 1)
    xc = 0.5 * (xs[1:] + xs[:-1])
    XC, YC, ZC = np.meshgrid(xc, xc, xc, indexing="ij")
    pe = pres(XC.transpose(2, 1, 0).ravel(), YC.transpose(2, 1, 0).ravel())
    return coord, mask, vals, ue, pe


def errors3(u, p, ue, pe):
    eu = np.linalg.norm(u.reshape(-1, 3) - ue) / np.linalg.norm(ue)
    pm, pem = p - p.mean(), pe - pe.mean()
    return eu, np.linalg.norm(pm - pem) / np.linalg.norm(pem)


def test_oracle_converges_to_closed_form_3d(oracle):
    import gale_amd as G
    errs = []
    for res in (4, 8):
        coord, mask, vals, ue, pe = setup3(G, res)
        o = oracle.Problem(3, (res, res, res), (0, 0, 0), (1, 1, 1), coord, np.zeros(len(coord), np.int32),
                           [dict(eta0=1.0, density=0.0)], bc_mask=mask, bc_value=vals, gravity=0.0)
        st, u, p, _ = o.uzawa(oracle.uzawa_params(tolerance=1e-10, max_iterations=5000, inner_rtol_override=1e-12,
                                                  inner_max_iterations=100000, nthreads=4))
        assert st == 0
        errs.append(errors3(u, p, ue, pe))
    print(errs)
    assert errs[0][0] / errs[1][0] > 3.3 and errs[0][1] / errs[1][1] > 1.8


@pytest.mark.gpu
def test_gpu_converges_to_closed_form_3d(gale):
    errs = []
    for res in (8, 16, 32):
        coord, mask, vals, ue, pe = setup3(gale, res)
        pb = gale.Problem(dim=3, res=(res, res, res), lo=(0, 0, 0), hi=(1, 1, 1), coord=coord,
                          material_index=np.zeros(len(coord), np.int32), materials=[gale.material(eta0=1.0)],
                          bc_mask=mask, bc_value=vals, gravity=0.0)
        s = gale.Stokes_SLE_UzawaSolver(pb, gale.uzawa_params(tolerance=1e-10, maxIterations=10000,
                                                             inner_rtol_override=1e-12))
        u, p, rep = s.Solve()
        assert rep.status == 0
        errs.append(errors3(u, p, ue, pe))
        print(f"res {res}^3: velocity {errs[-1][0]:.3e} pressure {errs[-1][1]:.3e}", flush=True)
        s.close()
        pb.close()
    for (eu0, ep0), (eu1, ep1) in zip(errs, errs[1:]):
        assert eu0 / eu1 > 3.5 and ep0 / ep1 > 1.8
    assert errs[-1][0] < 0.015 and errs[-1][1] < 0.15

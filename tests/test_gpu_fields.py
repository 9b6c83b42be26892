"""GPU output fields of the material points (gs_solver_particle_fields:
StoreStress / StoreViscosity / strain-rate invariant) against the oracle:
a linear velocity field (closed form), the solver's own solution of a tight
solve (2D sinker, 3D Rayleigh-Taylor) and a yielding Drucker-Prager Picard
solve."""
import numpy as np
import pytest

from conftest import rel
from test_gpu_parity import otight, pair, small_sinker, tight

pytestmark = pytest.mark.gpu


def _node_coords(oracle, cfg):
    dim = len(cfg.res)
    res, lo, hi = list(cfg.res), list(cfg.lo), list(cfg.hi)
    axes = [np.array([oracle.vertex_coord(res[d], lo[d], hi[d], i) for i in range(res[d] + 1)]) for d in range(dim)]
    grids = np.meshgrid(*axes, indexing="ij")
    return np.stack([g.transpose().ravel() for g in grids], axis=1)


@pytest.mark.parametrize("dim,res", [(2, 12), (3, 6)])
def test_fields_linear_velocity_and_solution(gale, oracle, dim, res):
    cfg = small_sinker(dim, res, ppc=6)
    g, o = pair(gale, oracle, cfg)
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale))
    u, p, rep = s.Solve()
    assert rep.status == 0
    st, ou, op, _ = o.uzawa(otight(oracle))
    assert st == 0
    # the solver's own solution
    f = s.particle_fields()
    fo = o.particle_fields(u)
    for k in ("strain_rate", "stress", "strain_rate_inv"):
        assert rel(f[k], fo[k]) < 1e-12, k
    assert np.array_equal(f["viscosity"], fo["viscosity"])
    # an explicit linear field: constant strain rate sym(A)
    A = np.random.default_rng(dim).normal(size=(dim, dim))
    ul = (_node_coords(oracle, cfg) @ A.T).ravel()
    fl = s.particle_fields(ul)
    S = 0.5 * (A + A.T)
    want = np.array([S[0, 0], S[1, 1], S[0, 1]]) if dim == 2 else \
        np.array([S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]])
    assert np.allclose(fl["strain_rate"], want, rtol=1e-11, atol=1e-11)
    assert np.allclose(fl["strain_rate_inv"], np.sqrt(0.5 * np.sum(S * S)), rtol=1e-11)
    assert np.allclose(fl["stress"], 2.0 * fl["viscosity"][:, None] * want, rtol=1e-11, atol=1e-11)
    s.close()
    g.close()


def test_fields_after_drucker_prager_picard(gale, oracle):
    """Yielded viscosities and stresses of a Picard solve (StoreVisc /
    StoreStress after the last constitutive pass)."""
    from gale_amd import configs
    cfg = configs.lithosphere_2d(nx=32, ny=8, ppc=10)
    g, o = pair(gale, oracle, cfg)
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale, tolerance=1e-11))
    u, p, rep = s.NonLinearExecute(nonLinearTolerance=1e-4, nonLinearMaxIterations=20)
    st, ou, op, orep = o.picard(otight(oracle, tolerance=1e-11), tolerance=1e-4, max_iterations=20)
    assert st == 0 and rep.status == 0
    f = s.particle_fields()
    fo = o.particle_fields(ou)
    assert rel(f["viscosity"], fo["viscosity"]) < 1e-6
    assert rel(f["stress"], fo["stress"]) < 1e-6
    assert rel(f["strain_rate_inv"], fo["strain_rate_inv"]) < 1e-6
    assert f["viscosity"].min() < 1e2 * (1 - 1e-6)  # the lid yields somewhere
    s.close()
    g.close()

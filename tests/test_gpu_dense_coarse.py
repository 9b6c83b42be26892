"""The dense coarsest multigrid solve on the 12^3 level (5577 free dofs with
free-slip walls): the inverse is formed on the device (Cholesky factor and
inverse) and applied by the wide row kernel.  A 48^3 hierarchy stops there
(48-24-12); with the bound lowered to 2000 free dofs it goes on to 6^3 and
the host-formed packed inverse.  Both V-cycles must give the same converged
solution, FP64 and mixed precision, and the exact 12^3 coarse solve must not
need more inner iterations."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
def test_dense_12_cubed_matches_deeper_hierarchy(gale, monkeypatch, prec):
    from gale_amd import configs
    cfg = configs.random_viscosity_3d(res=48, ppc=4)
    out = {}
    for bound in ("2000", "6000"):
        monkeypatch.setenv("GALE_MG_DENSE_MAX", bound)
        g = gale.Problem(**cfg.problem_kwargs())
        s = gale.Stokes_SLE_UzawaSolver(g, gale.uzawa_params(tolerance=1e-8, inner_rtol_override=1e-11,
                                                             precision=prec, maxIterations=5000))
        u, p, rep = s.Solve()
        out[bound] = (u, p, rep)
        print(f"bound {bound} prec {prec}: status {rep.status} uzawa {rep.uzawa_iterations} inner "
              f"{rep.inner_iterations} setup {rep.setup_seconds:.3f}s solve {rep.solve_seconds:.3f}s "
              f"rel {rep.rel_residual:.2e}", flush=True)
        assert rep.status == 0 and rep.rel_residual <= 1e-8
    (u0, p0, r0), (u1, p1, r1) = out["2000"], out["6000"]
    assert rel(u1, u0) < 1e-6 and rel(p1 - p1.mean(), p0 - p0.mean()) < 1e-5
    assert r1.inner_iterations <= 1.1 * r0.inner_iterations

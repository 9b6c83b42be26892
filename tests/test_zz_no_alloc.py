"""The no-allocation K apply (gs_apply_K_noalloc) through the C ABI: it
equals gs_apply_K bit for bit on two calls in a row, the second on a smaller
mesh, and refuses a mesh whose scratch would not fit its static array."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(gale, cfg, seed):
    import torch
    g = gale.Problem(**cfg.problem_kwargs())
    s = gale.Stokes_SLE_UzawaSolver(g)
    s.setup()
    rng = np.random.default_rng(seed)
    x = torch.tensor(rng.normal(size=g.N * g.dim), device="cuda:0")
    y_ref = torch.zeros_like(x)
    y_na = torch.full_like(x, np.nan)
    torch.cuda.synchronize()
    s.apply_K(x, y_ref)
    s.apply_K_noalloc(x, y_na)
    s.sync()
    a, b = y_ref.cpu().numpy(), y_na.cpu().numpy()
    s.close()
    g.close()
    return a, b


def test_noalloc_apply_equals_allocating_apply(gale):
    from gale_amd import configs
    print("no-alloc apply: 3D 24^3 then 3D 10^3 then 2D 20^2", flush=True)
    big = configs.random_viscosity_3d(res=24, ppc=6)
    small = configs.rayleigh_taylor_3d(res=10, ppc=5)
    flat = configs.sinker_2d(res=20, ppc=6)
    for cfg, seed in ((big, 1), (small, 2), (flat, 3)):
        a, b = _case(gale, cfg, seed)
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        np.testing.assert_array_equal(a, b)


def test_noalloc_apply_refuses_what_does_not_fit(gale):
    import torch
    res = (128, 128, 64)   # 8192 8x8x4 tiles: 80 MB of scratch > 64 MiB
    coord = gale.layout_random(3, res, (0, 0, 0), (1, 1, 1), 1, 5)
    g = gale.Problem(3, res, (0, 0, 0), (1, 1, 1), coord, np.zeros(len(coord), np.int32), [gale.material()])
    s = gale.Stokes_SLE_UzawaSolver(g)
    x = torch.zeros(g.N * 3, dtype=torch.float64, device="cuda:0")
    y = torch.zeros_like(x)
    with pytest.raises(gale.GaleError) as e:
        s.apply_K_noalloc(x, y)
    assert e.value.status == gale.E_INVALID
    # and one that is not set up is refused rather than set up (which allocates)
    cfg_small = (6, 6, 6)
    c2 = gale.layout_random(3, cfg_small, (0, 0, 0), (1, 1, 1), 2, 5)
    g2 = gale.Problem(3, cfg_small, (0, 0, 0), (1, 1, 1), c2, np.zeros(len(c2), np.int32), [gale.material()])
    s2 = gale.Stokes_SLE_UzawaSolver(g2)
    x2 = torch.zeros(g2.N * 3, dtype=torch.float64, device="cuda:0")
    with pytest.raises(gale.GaleError):
        s2.apply_K_noalloc(x2, torch.zeros_like(x2))
    s.close(); s2.close(); g.close(); g2.close()

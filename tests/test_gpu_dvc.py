"""GPU checks of the discrete Voronoi cell (DVC) integration weights
(gs_dvc_cells, gs_problem_set_weights) against the oracle: sub-cell
ownership is integer work, so centroids and weights must equal the oracle's
bit for bit; the reference's fixtures and PCDVC statistics follow."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel
from test_dvc_oracle import pcdvc_samples, pcdvc_statistics

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("key", ["2d", "3d"])
def test_dvc_reference_fixture(gale, oracle, key):
    f = json.load(open(os.path.join(GOLDEN, "dvc_weights.json")))[key]
    dim = len(f["res"])
    xi = np.array(f["xi"])
    cent, w = gale.dvc_cells(dim, f["res"], [0, len(xi)], xi)
    oc, ow = oracle.dvc_cell(dim, f["res"], xi)
    assert np.array_equal(cent, oc) and np.array_equal(w, ow)
    assert np.all((cent - np.array(f["centroids"])) ** 2 < f["tol_sq"])
    assert np.all((w - np.array(f["weights"])) ** 2 < f["tol_sq"])


def test_dvc_random_cells_bit_exact(gale, oracle):
    """Many cells of uneven particle counts (1..60), 2D and 3D, several
    sub-grids: every centroid and weight equals the oracle's."""
    rng = np.random.default_rng(11)
    for dim, res in ((2, (10, 10)), (2, (7, 13)), (3, (10, 10, 10)), (3, (6, 9, 12))):
        counts = rng.integers(1, 61, size=300)
        counts[:3] = (1, 2, 60)
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        xi = rng.uniform(-1.0, 1.0, (off[-1], dim))
        xi[:5] = np.clip(np.round(xi[:5] * 5) / 5, -1, 1)  # on sub-cell faces and the box boundary
        print(f"dim {dim} res {res}: {len(counts)} cells, {off[-1]} particles", flush=True)
        cent, w = gale.dvc_cells(dim, res, off, xi)
        for e in range(len(counts)):
            sl = slice(off[e], off[e + 1])
            oc, ow = oracle.dvc_cell(dim, res, xi[sl])
            assert np.array_equal(cent[sl], oc) and np.array_equal(w[sl], ow), (dim, res, e)
        vol = 2.0 ** dim
        sums = np.add.reduceat(w, off[:-1])
        assert np.allclose(sums, vol, atol=1e-12)


def test_pcdvc_statistics_on_gpu(gale, oracle):
    """The 1000 PCDVC test layouts in one call: bit-exact against the oracle
    and the reference's expected error statistics."""
    exp = json.load(open(os.path.join(GOLDEN, "pcdvc_expected.json")))
    samples = pcdvc_samples(gale.layout_random)
    ref = pcdvc_samples(oracle.layout_random)
    assert all(np.array_equal(a, b) for a, b in zip(samples, ref))
    xi = np.concatenate(samples)
    off = np.arange(0, 20 * len(samples) + 1, 20, dtype=np.int32)
    cent, w = gale.dvc_cells(2, (10, 10), off, xi)
    cents = [cent[off[i]:off[i + 1]] for i in range(len(samples))]
    ws = [w[off[i]:off[i + 1]] for i in range(len(samples))]
    for i in (0, 1, 499, 999):
        oc, ow = oracle.dvc_cell(2, (10, 10), samples[i])
        assert np.array_equal(cents[i], oc) and np.array_equal(ws[i], ow)
    for name, (mean, std) in pcdvc_statistics(cents, ws).items():
        e = exp[name]
        assert abs(mean - e["mean"]) < e["mean_tol"] and abs(std - e["std"]) < e["std_tol"], (name, mean, std)


@pytest.mark.parametrize("dim,res", [(2, 12), (3, 6)])
def test_problem_dvc_weights_solve_parity(gale, oracle, dim, res):
    """A problem switched to DVC weights: the integration points and weights
    equal the oracle's, the Stokes solve matches it at tight tolerance, and
    after one RK2 step both recompute the same weights."""
    from test_gpu_parity import otight, pair, small_sinker, tight
    cfg = small_sinker(dim, res, ppc=10)
    g, o = pair(gale, oracle, cfg)
    g.set_weights(gale.WEIGHTS_DVC, 10)
    o.set_weights(1, (10, 10, 10))
    _, gown, gxi, gw = g.particles()
    oown, oxi, ow, _ = o.particles()
    assert np.array_equal(gown, oown)
    assert np.array_equal(gxi, oxi) and np.array_equal(gw, ow)
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale))
    u, p, rep = s.Solve()
    st, ou, op, _ = o.uzawa(otight(oracle))
    assert st == 0 and rep.status == 0
    assert rel(u, ou) < 1e-8 and rel(p - p.mean(), op - op.mean()) < 1e-8
    dt = s.compute_dt(1.0)
    assert s.advect_rk2(dt) == 0
    coord, gown, gxi, gw = g.particles()
    o.set_coords(coord)
    oown, oxi, ow, _ = o.particles()
    assert np.array_equal(gown, oown) and np.array_equal(gxi, oxi) and np.array_equal(gw, ow)
    s.close()
    g.close()


def _compare_particles(g, o):
    coord, gown, gxi, gw = g.particles()
    oown, oxi, ow, ocoord = o.particles()
    assert g.P == o.P
    assert np.array_equal(coord, ocoord)
    assert np.array_equal(gown, oown)
    assert np.array_equal(gxi, oxi) and np.array_equal(gw, ow)


def test_population_control_designed_cells(gale, oracle):
    from test_dvc_oracle import cells_layout, two_by_two_kwargs
    rng = np.random.default_rng(2)
    coord = cells_layout(rng, True)
    mat = np.zeros(len(coord), np.int32)
    kw = two_by_two_kwargs(coord, mat)
    g = gale.Problem(dim=2, res=kw["res"], lo=kw["lo"], hi=kw["hi"], coord=coord, material_index=mat,
                     materials=[gale.material(eta0=1.0, density=1.0), gale.material(eta0=2.0, density=2.0)])
    o = oracle.Problem(**kw)
    ga = g.population_control()
    oa = o.population_control()
    assert ga == oa and ga[0] >= 1 and ga[1] >= 2
    _compare_particles(g, o)


@pytest.mark.parametrize("dim,res", [(2, 12), (3, 5)])
def test_population_control_matches_oracle_and_solves(gale, oracle, dim, res):
    """Uneven particle density (coordinates squeezed towards the origin) and
    two materials: splits, deletions and interface cells all occur; the new
    particle sets in the reference's array order (deleted indices taken by the
    last particle), their integration points and weights (each cell's final
    table order) and a tight Stokes solve equal the oracle's, and a second
    pass agrees too."""
    from test_gpu_parity import otight, small_sinker, tight
    cfg = small_sinker(dim, res, ppc=12)
    kw, okw = _squeezed(cfg, dim, 1.5 if dim == 2 else 1.3)  # no empty cell, no firewall
    g = gale.Problem(**kw)
    o = oracle.Problem(**okw)
    for rep in range(2):
        ga = g.population_control()
        oa = o.population_control()
        print(f"pass {rep}: added {ga[0]} deleted {ga[1]}", flush=True)
        assert ga == oa and (rep > 0 or ga[1] > 0)
        _compare_particles(g, o)
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale))
    u, p, rep_ = s.Solve()
    st, ou, op, _ = o.uzawa(otight(oracle))
    assert st == 0 and rep_.status == 0
    assert rel(u, ou) < 1e-8 and rel(p - p.mean(), op - op.mean()) < 1e-8
    s.close()
    g.close()


@pytest.mark.parametrize("dim,res,avg", [(2, 12, 20), (3, 5, 12)])
def test_spacefiller_pcdvc_problem_parity(gale, oracle, dim, res, avg):
    """The reference's PIC input pattern: a SpaceFiller (Sobol) layout over
    the whole box, DVC weights and a population-control pass.  The layout
    equals the oracle's bit for bit, as do the particles after population
    control, and the Stokes solve matches the oracle's at tight tolerance."""
    from gale_amd import configs
    from test_gpu_parity import otight, small_sinker, tight
    from test_sobol import GENS
    cfg = small_sinker(dim, res, ppc=4)
    kw, okw = cfg.problem_kwargs(), cfg.oracle_kwargs()
    coord = gale.layout_spacefiller(dim, cfg.res, cfg.lo, cfg.hi, avg_per_cell=avg)
    assert np.array_equal(coord, oracle.layout_spacefiller(dim, cfg.res, cfg.lo, cfg.hi, GENS[:dim],
                                                           avg_per_cell=avg))
    shapes = [configs.everywhere(0), configs.sphere(1, (0.5, 0.6, 0.5), 0.2)]
    mat = configs.shape_layout(dim, coord, shapes)
    kw.update(coord=coord, material_index=mat)
    okw.update(coord=coord, material=mat)
    kw.pop("eta0_particle", None)
    okw.pop("eta0_particle", None)
    g = gale.Problem(**kw)
    o = oracle.Problem(**okw)
    g.set_weights(gale.WEIGHTS_DVC, 10)
    o.set_weights(1, (10, 10, 10))
    ga = g.population_control()
    oa = o.population_control()
    print(f"population control: added {ga[0]} deleted {ga[1]}", flush=True)
    assert ga == oa
    _compare_particles(g, o)
    s = gale.Stokes_SLE_UzawaSolver(g, tight(gale))
    u, p, rep = s.Solve()
    st, ou, op, _ = o.uzawa(otight(oracle))
    assert st == 0 and rep.status == 0
    assert rel(u, ou) < 1e-8 and rel(p - p.mean(), op - op.mean()) < 1e-8
    s.close()
    g.close()


def _squeezed(cfg, dim, power=1.7):
    """The config's particles squeezed towards the origin: dense and sparse cells."""
    kw, okw = cfg.problem_kwargs(), cfg.oracle_kwargs()
    c = np.array(kw["coord"], dtype=np.float64)
    lo, hi = np.array(cfg.lo, float), np.array(cfg.hi, float)
    for d in range(dim):
        t = (c[:, d] - lo[d]) / (hi[d] - lo[d])
        c[:, d] = lo[d] + (hi[d] - lo[d]) * t ** power
    kw["coord"] = c
    okw["coord"] = c
    return kw, okw


@pytest.mark.parametrize("dim,res,upT", [(2, 12, 25.0), (2, 12, 6.0), (3, 5, 25.0), (3, 7, 4.0)])
def test_inflow_matches_oracle(gale, oracle, dim, res, upT):
    """The Inflow branch (PCDVC.cxx:769-1030 in 3D, 1204-1460 in 2D) on
    uneven layouts: sparse cells below Threshold x particlesPerCell and cells
    whose last particle sits far from its centroid are re-populated from the
    C library's rand() stream and thinned at random table slots; with upperT
    lowered, re-populated cells split too (random owned sub-cells), so the
    device's windows over re-populated cells meet cells whose split draws move
    the later offsets.  Two passes continue one stream; particle sets, caller
    order, integration points, weights and rand() calls equal the oracle's
    bit for bit, also for a pass entered mid-stream on a fresh problem."""
    from test_gpu_parity import small_sinker
    cfg = small_sinker(dim, res, ppc=12)
    kw, okw = _squeezed(cfg, dim)
    g = gale.Problem(**kw)
    o = oracle.Problem(**okw)
    prm = dict(upperT=upT, Inflow=True, particlesPerCell=12, Threshold=0.8, CentPosRatio=0.01, rand_seed=5)
    skip = 0
    for rep in range(2):
        ga = g.population_control(gale.pcdvc_params(**prm, rand_skip=skip))
        oa = o.population_control(**prm, rand_skip=skip)
        print(f"pass {rep}: added {ga[0]} deleted {ga[1]} draws {g.rand_draws}", flush=True)
        assert ga == oa and g.rand_draws == o.rand_draws > 0
        _compare_particles(g, o)
        skip += g.rand_draws
    # a pass entered mid-stream on a fresh problem (the stream replayed)
    g3 = gale.Problem(**kw)
    o3 = oracle.Problem(**okw)
    ga3 = g3.population_control(gale.pcdvc_params(**prm, rand_skip=7))
    oa3 = o3.population_control(**prm, rand_skip=7)
    assert ga3 == oa3 and g3.rand_draws == o3.rand_draws
    _compare_particles(g3, o3)
    g3.close()
    g.close()


@pytest.mark.parametrize("dim,res", [(2, 12), (3, 5)])
def test_population_control_firewall(gale, oracle, dim, res):
    """A denser squeeze leaves cells of 4 or 5 particles that delete down to
    2: the reference stops there (its under-resolved-cell firewall,
    PCDVC.cxx:1031-1033); the library and the oracle both refuse the pass."""
    from test_gpu_parity import small_sinker
    cfg = small_sinker(dim, res, ppc=12)
    kw, okw = _squeezed(cfg, dim, 1.7)
    g = gale.Problem(**kw)
    o = oracle.Problem(**okw)
    with pytest.raises(gale.GaleError):
        g.population_control()
    with pytest.raises(RuntimeError):
        o.population_control()
    g.close()

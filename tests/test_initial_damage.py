"""StrainWeakening's initial damage (host side, no GPU).

The reference draws it from the C library's rand() after srand(randomSeed)
(StrainWeakening.cxx:333-430).  The library restates that generator; it must
reproduce the C library's own rand() sequence (called here through ctypes)
for every seed.  The damage values must equal the oracle's, which draws from
the C library's srand/rand directly, bit for bit; with the whole fraction
damaged and no profiles the values are every second draw scaled by the final
softening strain (closed form)."""
import ctypes

import numpy as np
import pytest

LIBC = ctypes.CDLL("libc.so.6")
RAND_MAX = 2147483647


def libc_stream(seed, n):
    LIBC.srand(ctypes.c_uint(seed))
    return np.array([LIBC.rand() for _ in range(n)], np.int64)


@pytest.mark.parametrize("seed", [0, 1, 13, 4242, 2 ** 31 - 1, 2 ** 32 - 1])
def test_crand_matches_c_library(seed):
    import gale_amd
    assert np.array_equal(gale_amd.crand_stream(seed, 20000), libc_stream(seed, 20000))


def test_full_damage_closed_form():
    import gale_amd
    rng = np.random.default_rng(0)
    coord = np.c_[rng.uniform(size=(500, 2)), np.zeros(500)]
    w = gale_amd.initial_damage(2, coord, 1.0, 0.5, randomSeed=7)
    draws = libc_stream(7, 1000)
    assert np.array_equal(w, (1.0 * draws[1::2].astype(np.float64)) * 0.5 / RAND_MAX)
    assert np.all(gale_amd.initial_damage(2, coord, 0.0, 0.5, randomSeed=7) == 0.0)


@pytest.mark.parametrize("kw", [
    dict(initialDamageFraction=0.3, finalSofteningStrain=0.4, randomSeed=11),
    dict(initialDamageFraction=0.6, finalSofteningStrain=1.0, randomSeed=0, initialDamageFactor=0.5,
         initialDamageWavenumber=2.0),
    dict(initialDamageFraction=0.8, finalSofteningStrain=0.2, randomSeed=99, initialDamageWavenumberCosI=3.0,
         initialDamageWavenumberSinK=1.5, initialDamageWavenumberSinI=0.7, initialDamageWavenumberCosK=2.5),
])
@pytest.mark.parametrize("with_shape", [False, True])
def test_initial_damage_matches_oracle(oracle, kw, with_shape):
    import gale_amd
    rng = np.random.default_rng(3)
    coord = rng.uniform(size=(4000, 3)) * np.array([2.0, 1.0, 1.0])
    centre, radius = np.array([1.0, 0.5, 0.5]), 0.35
    shape = gale_amd.sphere(0, tuple(centre), radius) if with_shape else None
    w = gale_amd.initial_damage(3, coord, initialStrainShape=shape, **kw)
    inside = (np.sum((coord - centre) ** 2, axis=1) <= radius ** 2) if with_shape else None
    ow = oracle.initial_damage(coord, kw["initialDamageFraction"], kw["finalSofteningStrain"], kw["randomSeed"],
                               factor=kw.get("initialDamageFactor", 1.0),
                               wavenumber=kw.get("initialDamageWavenumber", -1.0),
                               sin_i=kw.get("initialDamageWavenumberSinI", -1.0),
                               cos_i=kw.get("initialDamageWavenumberCosI", -1.0),
                               sin_k=kw.get("initialDamageWavenumberSinK", -1.0),
                               cos_k=kw.get("initialDamageWavenumberCosK", -1.0), inside=inside)
    assert np.array_equal(w, ow)
    expect = kw["initialDamageFraction"] * (inside.mean() if with_shape else 1.0)
    assert 0.5 * expect < np.mean(w > 0) <= expect + 0.05

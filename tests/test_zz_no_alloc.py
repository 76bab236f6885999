"""pimc_density_noalloc (the reference's Density without a device allocation)
against pimc_density through the C ABI: two calls in a row on one GPU, the
second with a smaller input, so the fixed device array's leftover counts from
the first call must not show; and the size limit is refused."""
import numpy as np
import pytest

from pimc import _capi as A
from pimc.potentials import HarmonicTrap
from pimc.system import chain_desc

pytestmark = pytest.mark.gpu


def test_density_noalloc_matches_density():
    rng = np.random.default_rng(301)
    descs = [chain_desc(HarmonicTrap(1.0), dim=2, M=64, N=12, L=4.0, beta=1.0, lam=0.5, ell=8),
             chain_desc(HarmonicTrap(1.0), dim=2, M=16, N=3, L=4.0, beta=1.0, lam=0.5, ell=8)]
    opt = A.Options()
    opt.precision, opt.sweeps_per_launch, opt.seed = A.FP32, 4, 9
    ctx = A.Context(descs, opt)
    for c, d in enumerate(descs):
        ctx.set_paths(c, np.clip(rng.normal(0, 1.0, size=(d.N, 2, d.M)), -3.9, 3.9))
    ctx.run(8, measure=False)
    # first call: the larger chain and histogram
    want, nd_w = ctx.density(0, 200)
    got, nd_g = ctx.density_noalloc(0, 200)
    assert nd_g == nd_w == 64
    assert np.array_equal(got, want) and want.sum() > 0.9 * 12 * 64
    # second call: a smaller chain and histogram; the array still holds the first call's counts
    want, nd_w = ctx.density(1, 16)
    got, nd_g = ctx.density_noalloc(1, 16)
    assert nd_g == nd_w == 16
    assert np.array_equal(got, want) and 0 < want.sum() <= 3 * 16
    # accumulation into the caller's buffer, as pimc_density does
    acc, nd = ctx.density_noalloc(1, 16, dens=got, ndata=nd_g)
    assert nd == 32 and np.array_equal(acc, 2 * want)
    # the limit: 513 x 513 bins do not fit the fixed array
    with pytest.raises(A.PimcError) as e:
        ctx.density_noalloc(0, 513)
    assert "262144" in str(e.value)
    ctx.close()

"""GPU: the atoms' z order (qem_sortperm_z; src/QEM_long.jl:32-44, the
stable sortperm of z) is bit-exact against a stable sort by (z, index), and
against the oracle's, on uniform, layered, nearly-equal and monotone z, from
1 atom to above 2^20: one tile of the hand-written radix sort and tile
boundaries (1024 keys per tile up to 2^20 keys, more per tile beyond), with
eight digit passes over the 64-bit keys."""
import numpy as np
import pytest

from helpers import L, O
from zsort_cases import DISTS, LZ, stable_order, z_values

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 63, 1023, 1024, 1025, 4097, 10000, 65537, 100000, 131073, 1048577]


@pytest.fixture(scope="module")
def ctx():
    p = L.default_params()
    p.Lx, p.Ly, p.Lz = 100.0, 100.0, LZ
    p.r_cutoff, p.alpha, p.gamma_1, p.gamma_2 = 5.0, 0.5, 0.5, 0.5
    p.rbe_p = 0
    with L.Context(p) as c:
        yield c


@pytest.mark.parametrize("n", SIZES)
def test_z_order_all_distributions(ctx, n):
    rng = np.random.default_rng(n)
    for dist in DISTS:
        xyz = np.column_stack([rng.uniform(0, 100, n), rng.uniform(0, 100, n),
                               z_values(dist, n, seed=n)])
        got = ctx.sortperm_z(xyz)
        want = stable_order(xyz[:, 2])
        assert np.array_equal(got, want), (dist, n)
        if n <= 10000:
            assert np.array_equal(got, O.sortperm_z(xyz)), (dist, n)

"""Substrate_LJ wall (src/Substrate_LJ.jl:7-46): CPU pins for the oracle.

The oracle restates the reference as written.  It is checked here against
the wall potential U(dz) = 4 eps ((sigma/dz)^12 - (sigma/dz)^6) written out
independently: with the true-force option it must equal -dU/dz by central
differences, and as written it is that force divided by |dz| (the reference
applies the LJ "force over distance" factor as the force).  The cutoff
acts on the signed dz as written, so a wall with direction -1 acts at every
distance; the option bit moves it to |dz|.
"""
import numpy as np
import pytest

import oracle as O


def U(dz, sigma, eps):
    sr6 = (sigma / dz) ** 6
    return 4.0 * eps * (sr6 * sr6 - sr6)


@pytest.mark.parametrize("direction,sub_pos", [(+1, 0.0), (-1, 10.0)])
def test_true_force_is_minus_potential_gradient(direction, sub_pos):
    rng = np.random.default_rng(0)
    sigma, eps, cutoff = 0.5, 1.3, 3.0
    dz = direction * rng.uniform(0.35, 2.5, 200)
    z = sub_pos + dz
    xyz = np.column_stack([rng.uniform(0, 5, 200), rng.uniform(0, 5, 200), z])
    f, st = O.substrate_lj(xyz, cutoff, sigma, eps, sub_pos, direction, sem=3)
    assert st == 0
    h = 1e-6
    ref = -(U(dz + h, sigma, eps) - U(dz - h, sigma, eps)) / (2 * h)
    assert np.allclose(f[:, 2], ref, rtol=1e-7, atol=1e-9)
    assert np.all(f[:, :2] == 0.0)
    # as written: the same divided by |dz|
    fa, _ = O.substrate_lj(xyz, cutoff, sigma, eps, sub_pos, direction, sem=1)
    assert np.allclose(fa[:, 2] * np.abs(dz), f[:, 2], rtol=1e-13)


def test_as_written_cutoff_is_on_signed_dz():
    sigma, eps = 0.5, 1.0
    lower = np.array([[1.0, 1.0, 0.9], [1.0, 1.0, 1.1]])      # dz = 0.9, 1.1
    f, st = O.substrate_lj(lower, 1.0, sigma, eps, 0.0, +1)
    assert st == 0 and f[0, 2] != 0.0 and f[1, 2] == 0.0
    upper = np.array([[1.0, 1.0, 9.1], [1.0, 1.0, 2.0]])      # dz = -0.9, -8
    f, _ = O.substrate_lj(upper, 1.0, sigma, eps, 10.0, -1)
    assert f[0, 2] != 0.0 and f[1, 2] != 0.0                  # every distance
    assert f[1, 2] == pytest.approx(-24.0 * eps / 64.0 * (2 * (sigma / 8) ** 12 - (sigma / 8) ** 6),
                                    rel=1e-14)
    f, _ = O.substrate_lj(upper, 1.0, sigma, eps, 10.0, -1, sem=1)
    assert f[0, 2] != 0.0 and f[1, 2] == 0.0
    # repulsive inside 2^(1/6) sigma: away from each wall
    near = O.substrate_lj(np.array([[0, 0, 0.4]]), 1.0, sigma, eps, 0.0, +1)[0]
    assert near[0, 2] > 0
    near = O.substrate_lj(np.array([[0, 0, 9.6]]), 1.0, sigma, eps, 10.0, -1)[0]
    assert near[0, 2] < 0


def test_assertion_of_the_wall_side():
    xyz = np.array([[0.0, 0.0, 0.5], [0.0, 0.0, -0.1]])
    f, st = O.substrate_lj(xyz, 1.0, 0.5, 1.0, 0.0, +1)
    assert st == 1 and np.all(f == 0.0)
    f, st = O.substrate_lj(np.array([[0.0, 0.0, 10.0]]), 1.0, 0.5, 1.0, 10.0, -1)
    assert st == 1                                             # dz = 0 fails too
